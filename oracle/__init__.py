"""Oracle: CPU restatement of COFFE's spherical-Bessel integral path.

Test infrastructure only -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product (coffe_amd).
"""
from .oracle import *  # noqa: F401,F403
