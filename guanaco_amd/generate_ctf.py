"""python -m guanaco_amd.generate_ctf: the ``guanaco.generate_ctf`` command."""
from guanaco_amd.command_line import generate_ctf

if __name__ == "__main__":
    generate_ctf()
