"""The fold engine's options beyond the phase model, against the oracle's
restatement of Fold::fold (Fold.C:662-973): WeightedTimeSeries weights (a
zero weight leaves its samples unfolded and uncounted), zeroed data (hits per
channel, or in one row, counted where value 0 of pol 0 is non-zero, and the
number of such samples), and a fold replanned in place for the next block."""
import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _data(rng, nchan, npol, ndim, ndat):
    return rng.standard_normal((nchan, npol, ndat * ndim)).astype(np.float32)


@pytest.mark.parametrize("ndpw,widat", [(1000, 0), (4096, 123), (1, 0)])
def test_fold_weights(ctx, dspx, oracle, ndpw, widat):
    import torch
    rng = np.random.default_rng(ndpw + widat)
    nchan, npol, ndim, ndat, nbin = 2, 2, 1, 120_000, 257
    pps = 3.3e-4
    data = _data(rng, nchan, npol, ndim, ndat)
    nweights = (ndat + widat + ndpw - 1) // ndpw
    weights = (rng.uniform(size=nweights) > 0.3).astype(np.uint16) * rng.integers(1, 100, nweights).astype(np.uint16)
    chunks = [(0, 50_000, 0.17, 0), (55_000, 60_000, -1.4, 1)]
    prof_o = np.zeros((2, nchan, npol, nbin * ndim), np.float32)
    hits_o = np.zeros((2, 1, nbin), np.uint32)
    nf_o = 0
    plans = []
    for (i0, n, phi, isub) in chunks:
        _, _, nf, bp = oracle.fold(data, ndim, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof_o[isub],
                                   hits=hits_o[isub], weights=weights, ndatperweight=ndpw, weight_idat=widat,
                                   want_binplan=True)
        nf_o += nf
        plans.append((i0, n, isub, bp))
    print("fold weights", ndpw, widat, flush=True)
    f = dspx.Fold(ctx, ndat, pps, nbin, chunks, weights=weights, ndatperweight=ndpw, weight_idat=widat)
    assert f.ndat_folded() == nf_o
    plan = torch.zeros(ndat, dtype=torch.int32, device="cuda")
    f.binplan(plan)
    plan_h = from_dev(plan).view(np.uint32)
    for (i0, n, isub, bp) in plans:
        got = plan_h[i0:i0 + n]
        bad = bp == nbin
        assert bad.any() and not bad.all()
        assert np.all(got[bad] == 0xffffffff)
        assert np.array_equal(got[~bad] & 0xffff, bp[~bad])
        assert np.all((got[~bad] >> 16) == isub)
    hits = torch.zeros((2, nbin), dtype=torch.int32, device="cuda")
    f.hits(hits)
    assert np.array_equal(from_dev(hits).view(np.uint32), hits_o[:, 0])
    prof = torch.zeros((2, nchan, npol, nbin * ndim), dtype=torch.float32, device="cuda")
    f.execute(to_dev(data), npol * ndat * ndim, ndat * ndim, nchan, npol, ndim, prof)
    assert rel_rms(from_dev(prof), prof_o) < 1e-6
    f.close()


@pytest.mark.parametrize("per_channel", [True, False])
def test_fold_zeroed(ctx, dspx, oracle, per_channel):
    import torch
    rng = np.random.default_rng(7 + per_channel)
    nchan, npol, ndim, ndat, nbin = 3, 2, 4, 90_000, 128
    pps = 1.1e-3
    data = _data(rng, nchan, npol, ndim, ndat)
    # blocks of 1024 samples zapped per channel (both pols), as SK leaves them
    nblk = ndat // 1024
    zap = rng.uniform(size=(nchan, nblk)) < 0.2
    for c in range(nchan):
        for b in np.nonzero(zap[c])[0]:
            data[c, :, b * 1024 * ndim:(b + 1) * 1024 * ndim] = 0.0
    hits_nchan = nchan if per_channel else 1
    chunks = [(0, 40_000, 0.5, 0), (40_000, 45_000, 0.25, 1)]
    prof_o = np.zeros((2, nchan, npol, nbin * ndim), np.float32)
    hits_o = np.zeros((2, hits_nchan, nbin), np.uint32)
    nf_o = 0
    for (i0, n, phi, isub) in chunks:
        _, _, nf = oracle.fold(data, ndim, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof_o[isub],
                               hits=hits_o[isub], zeroed=True, hits_nchan=hits_nchan)
        nf_o += nf
    print("fold zeroed", per_channel, flush=True)
    # weights are ignored with zeroed data (Fold.C:728, 802)
    f = dspx.Fold(ctx, ndat, pps, nbin, chunks, weights=np.zeros(ndat // 100 + 1, np.uint16), ndatperweight=100,
                  zeroed=True)
    assert f.ndat_folded() == 0
    prof = torch.zeros((2, nchan, npol, nbin * ndim), dtype=torch.float32, device="cuda")
    hits = torch.zeros((2, hits_nchan, nbin), dtype=torch.int32, device="cuda")
    nfolded = torch.zeros(1, dtype=torch.int64, device="cuda")
    f.execute_zeroed(to_dev(data), npol * ndat * ndim, ndat * ndim, nchan, npol, ndim, prof, hits, hits_nchan,
                     nfolded)
    assert np.array_equal(from_dev(hits).view(np.uint32), hits_o)
    assert int(from_dev(nfolded)[0]) == nf_o
    assert rel_rms(from_dev(prof), prof_o) < 1e-6
    with pytest.raises(dspx.DspxError):
        f.hits(hits)  # zeroed data: the plan's hits are not the reference's
    f.close()


def test_fold_replan(ctx, dspx, oracle):
    """A fold replanned for a second block (other phases, chunks, bins and
    block length) equals a fold made for that block."""
    import torch
    rng = np.random.default_rng(11)
    f = dspx.Fold(ctx, 100_000, 2e-4, 512, [(0, 100_000, 0.3, 0)])
    for ndat, pps, nbin, chunks in [(70_000, 7e-5, 100, [(1000, 30_000, 0.9, 0), (31_000, 39_000, 0.1, 1)]),
                                    (130_000, 1e-3, 2048, [(0, 130_000, -0.6, 0)])]:
        data = _data(rng, 2, 1, 1, ndat)
        nsub = max(c[3] for c in chunks) + 1
        prof_o = np.zeros((nsub, 2, 1, nbin), np.float32)
        hits_o = np.zeros((nsub, 1, nbin), np.uint32)
        for (i0, n, phi, isub) in chunks:
            oracle.fold(data, 1, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof_o[isub], hits=hits_o[isub])
        print("fold replan", ndat, nbin, flush=True)
        f.replan(ndat, pps, nbin, chunks)
        assert f.ndat_folded() == sum(c[1] for c in chunks)
        hits = torch.zeros((nsub, nbin), dtype=torch.int32, device="cuda")
        f.hits(hits)
        assert np.array_equal(from_dev(hits).view(np.uint32), hits_o[:, 0])
        prof = torch.zeros((nsub, 2, 1, nbin), dtype=torch.float32, device="cuda")
        f.execute(to_dev(data), ndat, ndat, 2, 1, 1, prof)
        assert rel_rms(from_dev(prof), prof_o) < 1e-6
    f.close()


def test_engine_outlives_context(dspx, oracle):
    """dspx_ctx_destroy before an engine made on it: the engine keeps the
    context (and its stream) alive, runs, and frees it when destroyed."""
    import ctypes as C
    import torch
    L = dspx.lib()
    h = C.c_void_p()
    assert L.dspx_ctx_create(0, None, C.byref(h)) == 0  # the library's own stream
    ndat, nbin, pps = 50_000, 64, 1e-3
    rng = np.random.default_rng(2)
    data = rng.standard_normal((1, 1, ndat)).astype(np.float32)
    prof_o, hits_o, _ = oracle.fold(data, 1, 0.2, pps, nbin)
    arr = (dspx.FoldChunk * 1)(dspx.FoldChunk(0, ndat, 0.2, 0, 0))
    f = C.c_void_p()
    assert L.dspx_fold_create(h, ndat, pps, nbin, arr, 1, C.byref(f)) == 0
    d_in = to_dev(data)
    prof = torch.zeros(nbin, dtype=torch.float32, device="cuda")
    hits = torch.zeros(nbin, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    print("engine after its context", flush=True)
    L.dspx_ctx_destroy(h)  # the caller's reference only
    assert L.dspx_fold_hits(f, hits.data_ptr()) == 0
    assert L.dspx_fold_execute(f, d_in.data_ptr(), ndat, ndat, 1, 1, 1, prof.data_ptr()) == 0
    L.dspx_fold_destroy(f)  # waits for the context's stream, then frees it
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(hits).view(np.uint32), hits_o[0])
    assert rel_rms(from_dev(prof), prof_o.ravel()) < 1e-6
