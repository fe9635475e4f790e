"""CPU oracle for the CSR builder and pull PageRank (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker; never by the product package.  See oracle.h for the functions
and the reference lines each one restates.
"""
from .oracle import *  # noqa: F401,F403
