"""CPU oracle for the NS3D WENO5-RK4 path (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package; the product path never does.
"""
