"""Writes sk_limits.npz: the SK threshold tables the GPU tests use (counts
0..32 in full, count 64 of M = 64, count 256 of M = 256 and count 1024 of
M = 1024; for the adaptive-SK tests, every count up to M = 64, 128, 256, 512
and 1024), from
dspsr-kanyixia_amd/sklimits.py,
so that GPU test processes do not evaluate the Pearson integrals themselves.

    python tests/golden/make_sk_limits.py
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
S = importlib.import_module("dspsr-kanyixia_amd.sklimits")

out = {}
for M, counts in ((32, None), (64, [64]), (256, [256]), (1024, [1024])):
    lo, up, sym, pcsd = S.limit_tables(M, 3.0, npol=2, counts=counts)
    out["lower%d" % M], out["upper%d" % M], out["sym%d" % M] = lo, up, sym
    out["pcsd%d" % M] = np.float64(pcsd)
# the resolutions of the adaptive-SK tests (blocks of 64 to 1024 time
# scrunched from blocks of 64 or 128): every count up to M (detect_fscr
# looks up the rounded mean count)
for M in (64, 128, 256, 512, 1024):
    lo, up, sym, pcsd = S.limit_tables(M, 3.0, npol=2)
    out["rlower%d" % M], out["rupper%d" % M], out["rsym%d" % M] = lo, up, sym
np.savez(os.path.join(HERE, "sk_limits.npz"), **out)
print("wrote", os.path.join(HERE, "sk_limits.npz"))
