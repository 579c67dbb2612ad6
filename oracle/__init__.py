"""TEST INFRASTRUCTURE: the FP64 CPU oracle that restates the reference substep.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
