"""Test infrastructure: the CPU oracle of the 2D TE path (see te_numpy.py).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker; the etalon_amd package never does.
"""
