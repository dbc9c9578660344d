"""python -m guanaco_amd.correct: the ``guanaco.correct`` command."""
from guanaco_amd.command_line import correct

if __name__ == "__main__":
    correct()
