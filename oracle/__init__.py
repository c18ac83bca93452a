"""Oracle: CPU restatement of cantucci's distance-estimation path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker (or the timed reference
CPU path), never by the product.  See oracle.cpp for the pins and the
functions that stay "parity unpinned".
"""
from .binding import *  # noqa: F401,F403
