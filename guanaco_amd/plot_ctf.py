"""python -m guanaco_amd.plot_ctf: the ``guanaco.plot_ctf`` command."""
from guanaco_amd.command_line import plot_ctf

if __name__ == "__main__":
    plot_ctf()
