"""python -m sgl_analytic_amd: the command line of cli.py."""
from .cli import main

main()
