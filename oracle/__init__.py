"""Test infrastructure: the CPU restatement of the reference's Lanczos path
(nv_oracle.c) and an independent dense brute force (dense.py).  Only tests,
the smoke check and bench.py's CPU baseline use this package."""
