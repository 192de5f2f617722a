"""CPU oracle for the D2Q9 lattice-Boltzmann step (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product never does.
"""
