"""ORACLE -- test infrastructure only.

CPU restatements of the reference algorithm used as the checker for the HIP
path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import anything under this package; the product never does.
"""
