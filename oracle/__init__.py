"""Test infrastructure: the FP64 oracle and the CPU baseline of the hot path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only to check results against.  The product package never imports it.
"""
