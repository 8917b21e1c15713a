"""`python -m pulsarsr_amd <command>`: the `psr` drop-in (see pulsarsr_amd/cli.py)."""
import sys

from .cli import main

main()
sys.exit(0)
