"""CPU oracle for the ExteriorPoissonRecon solve path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker.
"""
