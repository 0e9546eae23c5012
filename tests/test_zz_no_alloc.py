"""dspx_fold_static, the fold that allocates nothing, against the planned
engine (dspx_fold_create + dspx_fold_hits + dspx_fold_execute) on the same
inputs: two calls in a row through the C ABI, the second on a smaller block
with other chunks, so that the shared segment table is rewritten.  Hits
bit-exact; profiles within the fold's float tolerance (both accumulate with
atomics, in orders of their own)."""
import ctypes as C

import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _planned(L, h, ndat, pps, nbin, arr, n, d_in, nchan, npol, ndim, nsub):
    import torch
    f = C.c_void_p()
    assert L.dspx_fold_create(h, ndat, pps, nbin, arr, n, C.byref(f)) == 0
    hits = torch.zeros((nsub, nbin), dtype=torch.int32, device="cuda")
    prof = torch.zeros((nsub, nchan, npol, nbin * ndim), dtype=torch.float32, device="cuda")
    assert L.dspx_fold_hits(f, hits.data_ptr()) == 0
    assert L.dspx_fold_execute(f, d_in.data_ptr(), npol * ndat * ndim, ndat * ndim, nchan, npol, ndim,
                               prof.data_ptr()) == 0
    L.dspx_fold_destroy(f)
    return from_dev(hits).view(np.uint32), from_dev(prof)


def test_fold_static_matches_planned(ctx, dspx, oracle):
    import torch
    L = dspx.lib()
    rng = np.random.default_rng(5)
    cases = [
        # ndat, pps, nbin, ndim, chunks (sorted)
        (400_000, 1.3e-4, 1024, 2, [(0, 150_000, 0.7, 0), (150_000, 150_000, 0.2, 1), (310_000, 90_000, -0.3, 2)]),
        (90_000, 4.1e-3, 96, 1, [(5_000, 80_000, 0.05, 0)]),
    ]
    nchan, npol = 3, 2
    for ndat, pps, nbin, ndim, chunks in cases:
        nsub = max(c[3] for c in chunks) + 1
        data = rng.standard_normal((nchan, npol, ndat * ndim)).astype(np.float32)
        d_in = to_dev(data)
        arr = (dspx.FoldChunk * len(chunks))(*[dspx.FoldChunk(c[0], c[1], c[2], c[3], 0) for c in chunks])
        hits_p, prof_p = _planned(L, ctx.h, ndat, pps, nbin, arr, len(chunks), d_in, nchan, npol, ndim, nsub)
        hits = torch.zeros((nsub, nbin), dtype=torch.int32, device="cuda")
        prof = torch.zeros((nsub, nchan, npol, nbin * ndim), dtype=torch.float32, device="cuda")
        print("fold static", ndat, nbin, ndim, flush=True)
        rc = L.dspx_fold_static(ctx.h, ndat, pps, nbin, arr, len(chunks), d_in.data_ptr(), npol * ndat * ndim,
                                ndat * ndim, nchan, npol, ndim, prof.data_ptr(), hits.data_ptr())
        assert rc == 0, L.dspx_last_error(ctx.h)
        got_h = from_dev(hits).view(np.uint32)
        assert got_h.sum() == sum(c[1] for c in chunks)
        assert np.array_equal(got_h, hits_p)
        assert rel_rms(from_dev(prof), prof_p) < 1e-6
    # chunks out of order are refused (the call does not sort them)
    bad = (dspx.FoldChunk * 2)(dspx.FoldChunk(10, 5, 0.0, 0, 0), dspx.FoldChunk(0, 5, 0.0, 0, 0))
    assert L.dspx_fold_static(ctx.h, 20, 0.01, 8, bad, 2, None, 0, 0, 1, 1, 1, None, None) != 0
