"""Test infrastructure: CPU restatements of DASP's Shack-Hartmann centroiding path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import anything here,
and only as the checker (or the timed CPU baseline). The product package ``dasp_amd`` never does.
"""
