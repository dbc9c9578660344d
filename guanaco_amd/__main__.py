"""python -m guanaco_amd: the ``guanaco`` reconstruction command."""
from guanaco_amd.command_line import main

main()
