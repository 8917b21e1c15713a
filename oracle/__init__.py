"""The oracle: a CPU restatement of PulsarSr's search path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker.  The product (pulsarsr_amd) never uses it.
"""
from .pyoracle import *  # noqa: F401,F403
