"""The spectral-kurtosis engine on the GPU through the C ABI against the
oracle: block sums and counts (FPT float input and raw TFP 8/16-bit input),
the three detectors with and without omit_outliers, and the mask.  Counts,
masks and masked samples are bit-exact; S1/S2 (float results of double sums,
reduced in another order) agree to one float ulp."""
import os

import numpy as np
import pytest

from helpers import from_dev, sk_rfi_input, to_dev

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sk_limits.npz")


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.fixture(scope="module")
def tables():
    z = np.load(GOLDEN)
    return {M: (z["lower%d" % M], z["upper%d" % M], z["sym%d" % M], float(z["pcsd%d" % M])) for M in (32, 64)}


def _sk(dspx, ctx, tables, nchan, M, omit=False):
    sk = dspx.SpectralKurtosis(ctx, nchan, 2, M, omit_outliers=omit)
    sk.set_limits(*tables[M])
    return sk


@pytest.mark.parametrize("zeroed", [False, True])
def test_sk_compute_fpt(dspx, ctx, oracle, tables, zeroed):
    import torch
    M, nchan, nblk = 32, 9, 70
    x = sk_rfi_input(nchan, 2, nblk, M, seed=11, tone_chan=2, zero_blk=5)
    sums_o, counts_o = oracle.sk_compute(x, M, zeroed=zeroed)
    sk = _sk(dspx, ctx, tables, nchan, M)
    sums = torch.zeros((nblk, nchan, 2, 2), dtype=torch.float32, device="cuda")
    counts = torch.zeros((nblk, nchan), dtype=torch.int32, device="cuda")
    print("sk compute fpt zeroed", zeroed, flush=True)
    sk.compute(to_dev(x.view(np.float32)), 2 * x.shape[2] * 2, x.shape[2] * 2, nblk, sums, counts, zeroed=zeroed)
    assert np.array_equal(from_dev(counts).view(np.uint32), counts_o)
    u = _ulps(from_dev(sums), sums_o)
    assert u.max() <= 1 and (u > 0).mean() < 0.01


@pytest.mark.parametrize("nbit,nchan", [(8, 37), (16, 37), (8, 36)])
def test_sk_compute_raw(dspx, ctx, oracle, tables, nbit, nchan):
    """(8-bit samples of two pols with a multiple of 4 channels take the
    kernel that loads four channels' words at once.)"""
    import torch
    M, npol, nblk, t0 = 32, 2, 20, 13
    ndat = t0 + nblk * M + 5
    rng = np.random.default_rng(nbit)
    if nbit == 8:
        raw = rng.integers(-40, 41, (ndat, nchan, npol, 2)).astype(np.int8).view(np.uint8)
        x = oracle.unpack8(raw, nchan, npol, ndat, scale=0.25)
    else:
        raw = rng.integers(-3000, 3001, (ndat, nchan, npol, 2)).astype(np.int16)
        x = oracle.unpack16(raw, nchan, npol, ndat)
    xc = x.view(np.complex64).reshape(nchan, npol, ndat)[:, :, t0:]
    sums_o, counts_o = oracle.sk_compute(np.ascontiguousarray(xc[:, :, :nblk * M]), M)
    sk = _sk(dspx, ctx, tables, nchan, M)
    sums = torch.zeros((nblk, nchan, npol, 2), dtype=torch.float32, device="cuda")
    counts = torch.zeros((nblk, nchan), dtype=torch.int32, device="cuda")
    print("sk compute raw", nbit, flush=True)
    sk.compute_raw(to_dev(raw), nbit, t0, nblk, sums, counts, scale=0.25)
    assert np.array_equal(from_dev(counts).view(np.uint32), counts_o)
    u = _ulps(from_dev(sums), sums_o)
    assert u.max() <= 1 and (u > 0).mean() < 0.01


@pytest.mark.parametrize("flags,omit", [(1, False), (2, False), (4, False), (7, False), (3, True), (7, True)])
def test_sk_detect(dspx, ctx, oracle, tables, flags, omit):
    import torch
    M, nchan, nblk = 64, 16, 512
    lower, upper, sym, pcsd = tables[M]
    x = sk_rfi_input(nchan, 2, nblk, M, seed=3, tone_chan=5, impulse_blk=17)
    sums_o, counts_o = oracle.sk_compute(x, M)
    # detect_tscr's thresholds of the totals detect_skfb leaves below 32768
    # (omit_outliers), as the wrapper evaluates them for the device
    _, _, c_after = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=flags & 3,
                                     omit_outliers=omit)
    lim = {}
    if flags & 4:
        from importlib import import_module
        S = import_module("dspsr-kanyixia_amd.sklimits")
        for t in {int(t) for t in c_after.astype(np.uint64).sum(axis=0) if M < t < 32768}:
            L = S.SKLimits(t, 3.0, 1, 2)
            lim[t] = (L.lower, L.upper)
    mask_o, s2_o, c2_o = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=flags,
                                          omit_outliers=omit, tscr_limits=lim)
    assert mask_o.any()
    sk = dspx.SpectralKurtosis(ctx, nchan, 2, M, omit_outliers=omit)
    sk.set_limits(*tables[M], std_devs=3.0)
    sums, counts = to_dev(sums_o), to_dev(counts_o.view(np.int32))
    mask = torch.full((nblk, nchan), 7, dtype=torch.uint8, device="cuda")
    print("sk detect", flags, omit, flush=True)
    sk.detect(sums, counts, nblk, flags, mask)
    assert np.array_equal(from_dev(mask), mask_o)
    assert np.array_equal(from_dev(sums), s2_o) and np.array_equal(from_dev(counts).view(np.uint32), c2_o)


def test_sk_mask(dspx, ctx, oracle, tables):
    import torch
    M, nchan, nblk = 32, 5, 40
    rng = np.random.default_rng(1)
    x = sk_rfi_input(nchan, 2, nblk, M, seed=6)
    mask = (rng.random((nblk, nchan)) < 0.3).astype(np.uint8)
    expect = oracle.sk_mask(x, M, mask)
    sk = _sk(dspx, ctx, tables, nchan, M)
    n = x.shape[2]
    out = torch.full((nchan, 2, n, 2), 5.0, dtype=torch.float32, device="cuda")
    print("sk mask", flush=True)
    sk.mask(to_dev(mask), to_dev(x.view(np.float32)), 2 * n * 2, n * 2, out, 2 * n * 2, n * 2, nblk)
    assert np.array_equal(from_dev(out).view(np.complex64).reshape(nchan, 2, n), expect)


def test_sk_detect_requires_limits(dspx, ctx):
    import torch
    sk = dspx.SpectralKurtosis(ctx, 4, 2, 32)
    z = torch.zeros(64, dtype=torch.float32, device="cuda")
    with pytest.raises(dspx.DspxError) as e:
        sk.detect(z, z, 2, 1, z)
    assert e.value.code == dspx.E_STATE


@pytest.mark.parametrize("flags,omit", [(4, False), (7, True)])
def test_sk_tscr_every_total(dspx, ctx, oracle, tables, flags, omit):
    """detect_tscr on channel totals between M and 32768, which differ from
    channel to channel (zeroed samples; with omit_outliers, the blocks that
    detect_skfb zaps): each total is tested with SKLimits(total) as
    get_limits does (SpectralKurtosis.C:989-1012, 1045), the host evaluating
    the totals the device reports.  Masks, sums and counts bit-exact."""
    import torch
    from importlib import import_module
    S = import_module("dspsr-kanyixia_amd.sklimits")
    M, nchan, nblk = 32, 12, 90
    lower, upper, sym, pcsd = tables[M]
    x = sk_rfi_input(nchan, 2, nblk, M, seed=21, tone_chan=7, zero_blk=3)
    # a weak tone in channel 2: no block of it passes detect_skfb's
    # thresholds, its total does pass detect_tscr's
    x[2] += (10 * np.exp(2j * np.pi * 0.31 * np.arange(x.shape[2]))).astype(np.complex64)
    sums_o, counts_o = oracle.sk_compute(x, M, zeroed=True)
    # the totals detect_tscr meets: after detect_skfb where it zeroes counts
    _, _, c_after = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=flags & 3,
                                     omit_outliers=omit)
    totals = sorted({int(t) for t in c_after.astype(np.uint64).sum(axis=0) if t > M})
    assert len(totals) >= 2 and all(M < t < 32768 for t in totals)
    lim = {}
    for t in totals:
        L = S.SKLimits(t, 3.0, 1, 2)
        lim[t] = (L.lower, L.upper)
    mask_o, s2_o, c2_o = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=flags,
                                          omit_outliers=omit, tscr_limits=lim)
    assert mask_o[:, 7].all() and mask_o[:, 2].all() and not mask_o.all()
    sk = dspx.SpectralKurtosis(ctx, nchan, 2, M, omit_outliers=omit)
    # without the totals' thresholds the detector refuses
    sk.set_limits(lower, upper, sym, pcsd)
    sums, counts = to_dev(sums_o), to_dev(counts_o.view(np.int32))
    mask = torch.full((nblk, nchan), 7, dtype=torch.uint8, device="cuda")
    print("sk tscr without thresholds", flags, omit, flush=True)
    with pytest.raises(dspx.DspxError) as e:
        sk.detect(sums, counts, nblk, flags, mask)
    assert e.value.code == dspx.E_STATE
    # the wrapper evaluates SKLimits for the totals the device reports
    sk.set_limits(lower, upper, sym, pcsd, std_devs=3.0)
    sums, counts = to_dev(sums_o), to_dev(counts_o.view(np.int32))
    print("sk tscr every total", flags, omit, totals, flush=True)
    sk.detect(sums, counts, nblk, flags, mask)
    assert sorted(sk._tscr) == totals
    assert np.array_equal(from_dev(mask), mask_o)
    assert np.array_equal(from_dev(sums), s2_o) and np.array_equal(from_dev(counts).view(np.uint32), c2_o)
    sk.close()


@pytest.mark.parametrize("nov", [2, 4])
def test_sk_noverlap(dspx, ctx, oracle, tables, nov):
    """Overlapping blocks (noverlap 2 and 4): sums and counts from FPT float
    and raw 8-bit input, all three detectors (detect_tscr over every
    noverlap-th block) and the mask's output against the oracle."""
    import torch
    M, nchan, npol = 32, 6, 2
    ndat = 32 * 70 + 9
    x = sk_rfi_input(nchan, npol, 71, M, seed=30 + nov, tone_chan=2)[:, :, :ndat]
    x = np.ascontiguousarray(x)
    sk = dspx.SpectralKurtosis(ctx, nchan, npol, M, noverlap=nov)
    sk.set_limits(*tables[M], std_devs=3.0)
    npart, nout = sk.blocks(ndat)
    assert (npart, nout) == oracle.sk_blocks(ndat, M, nov)
    sums_o, counts_o = oracle.sk_compute(x, M, noverlap=nov)
    sums = torch.zeros((npart, nchan, npol, 2), dtype=torch.float32, device="cuda")
    counts = torch.zeros((npart, nchan), dtype=torch.int32, device="cuda")
    print("sk noverlap", nov, "compute", flush=True)
    sk.compute(to_dev(x.view(np.float32)), npol * ndat * 2, ndat * 2, npart, sums, counts)
    assert np.array_equal(from_dev(counts).view(np.uint32), counts_o)
    u = _ulps(from_dev(sums), sums_o)
    assert u.max() <= 1
    # raw 8-bit TFP input from sample t0
    rng = np.random.default_rng(nov)
    t0 = 5
    raw = rng.integers(-40, 41, (ndat + t0, nchan, npol, 2)).astype(np.int8).view(np.uint8)
    xr = oracle.unpack8(raw, nchan, npol, ndat + t0).view(np.complex64).reshape(nchan, npol, ndat + t0)
    sr_o, cr_o = oracle.sk_compute(np.ascontiguousarray(xr[:, :, t0:]), M, noverlap=nov)
    sk.compute_raw(to_dev(raw), 8, t0, npart, sums, counts)
    assert np.array_equal(from_dev(counts).view(np.uint32), cr_o)
    assert _ulps(from_dev(sums), sr_o).max() <= 1
    # detectors on the oracle's sums (tscr's totals below 32768: SKLimits of each)
    from importlib import import_module
    S = import_module("dspsr-kanyixia_amd.sklimits")
    lower, upper, sym, pcsd = tables[M]
    tot = {int(t) for t in counts_o[::nov].astype(np.uint64).sum(axis=0)}
    lim = {t: (S.SKLimits(t, 3.0, 1, 2).lower, S.SKLimits(t, 3.0, 1, 2).upper) for t in tot if M < t < 32768}
    mask_o, _, _ = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=7, tscr_limits=lim,
                                    noverlap=nov)
    assert mask_o[:, 2].all() and not mask_o.all()
    mask = torch.zeros((npart, nchan), dtype=torch.uint8, device="cuda")
    sk.detect(to_dev(sums_o), to_dev(counts_o.view(np.int32)), npart, 7, mask)
    assert np.array_equal(from_dev(mask), mask_o)
    # the mask's output: input samples M - M/noverlap on, block by block
    out_o = oracle.sk_mask(x, M, mask_o, noverlap=nov)
    out = torch.full((nchan, npol, ndat, 2), 3.0, dtype=torch.float32, device="cuda")
    sk.mask(mask, to_dev(x.view(np.float32)), npol * ndat * 2, ndat * 2, out, npol * ndat * 2, ndat * 2, npart)
    got = from_dev(out).view(np.complex64).reshape(nchan, npol, ndat)
    assert np.array_equal(got[:, :, :nout], out_o[:, :, :nout])
    sk.close()


def test_sk_channel_range(dspx, ctx, oracle, tables):
    """Excluded channels: out of detect_fscr's average and detect_tscr's
    tests, still tested by detect_skfb."""
    import torch
    M, nchan, nblk = 64, 16, 512
    lower, upper, sym, pcsd = tables[M]
    x = sk_rfi_input(nchan, 2, nblk, M, seed=3, tone_chan=5)
    sums_o, counts_o = oracle.sk_compute(x, M)
    inc = np.ones(nchan, np.uint8)
    inc[5] = 0
    inc[11:14] = 0
    mask_o, _, _ = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=7, channels=inc)
    sk = _sk(dspx, ctx, tables, nchan, M)
    sk.set_channels(inc)
    mask = torch.zeros((nblk, nchan), dtype=torch.uint8, device="cuda")
    print("sk channel range", flush=True)
    sk.detect(to_dev(sums_o), to_dev(counts_o.view(np.int32)), nblk, 7, mask)
    assert np.array_equal(from_dev(mask), mask_o)
    sk.set_channels(None)
    mask_all, _, _ = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=7)
    sk.detect(to_dev(sums_o), to_dev(counts_o.view(np.int32)), nblk, 7, mask)
    assert np.array_equal(from_dev(mask), mask_all) and not np.array_equal(mask_all, mask_o)
    sk.close()


RES_SPECS = {"64-256-1024": [(64, 1), (256, 4), (1024, 16)], "128-512": [(128, 2), (512, 8)]}


@pytest.fixture(scope="module")
def rtables():
    z = np.load(GOLDEN)
    return {M: (z["rlower%d" % M], z["rupper%d" % M], z["rsym%d" % M]) for M in (64, 128, 256, 512, 1024)}


@pytest.mark.parametrize("omit", [False, True])
@pytest.mark.parametrize("flags", [1, 2, 3, 7])
@pytest.mark.parametrize("spec", list(RES_SPECS))
def test_sk_resolutions(dspx, ctx, oracle, rtables, spec, flags, omit):
    """Adaptive SK: the detectors of every resolution over the finest
    resolution's sums (the oracle's, so that the comparison is of the
    detection), masks bit-exact with the oracle's; a tone, a burst and a
    channel whose power steps every 256 samples (seen only by the coarser
    blocks); the last resolution's time-scrunched test with SKLimits of the
    totals it meets; channel ranges on the 64-256-1024 case."""
    import torch
    res = RES_SPECS[spec]
    M0, nov0 = res[0]
    nchan, nblk = 24, 96
    x = sk_rfi_input(nchan, 2, nblk * 64 // M0, M0, seed=31, tone_chan=5, impulse_blk=3)
    n = x.shape[2]
    x[11] *= np.where((np.arange(n) // 256) % 2 == 0, 1.0, 2.5).astype(np.float32)
    sums_o, counts_o = oracle.sk_compute(x, M0, noverlap=nov0)
    np0 = sums_o.shape[0]
    include = None
    if spec == "64-256-1024":
        include = np.ones(nchan, np.uint8)
        include[18:] = 0
    sk = dspx.SpectralKurtosis(ctx, nchan, 2, M0, omit_outliers=omit, noverlap=nov0)
    lo, up, sy = rtables[M0]
    z = np.load(GOLDEN)
    sk.set_limits(lo, up, sy, float(z["pcsd1024"]), std_devs=3.0)
    if include is not None:
        sk.set_channels(include)
    for M, nov in res[1:]:
        sk.add_resolution(M, nov, *rtables[M], include=include, std_devs=3.0)
    assert sk.blocks_res(n)[0] == np0
    sums, counts = to_dev(sums_o), to_dev(counts_o.view(np.int32))
    mask = torch.zeros((np0, nchan), dtype=torch.uint8, device="cuda")
    print("sk resolutions", spec, flags, omit, flush=True)
    sk.detect_res(sums, counts, n, flags, mask)
    tl = dict(sk._tscr)
    m_o, s_o, c_o = oracle.sk_detect_res(sums_o, counts_o, n,
                                         [(M, nov) + rtables[M] + (include,) for M, nov in res],
                                         flags=flags, omit_outliers=omit, pcsd=float(z["pcsd1024"]),
                                         tscr_limits=tl)
    got = from_dev(mask)
    print("rows flagged %d, oracle %d" % (got.sum(), m_o.sum()), flush=True)
    assert np.array_equal(got, m_o)
    assert 0 < m_o.mean() < 1
    # the finest resolution's sums and counts as the oracle leaves its first rows
    if len(res) == 1:
        assert np.array_equal(from_dev(counts).view(np.uint32), c_o)
    sk.close()


def test_sk_mask_with_resolutions(dspx, ctx, oracle, rtables):
    import torch
    M0, nchan, nblk = 64, 6, 50
    x = sk_rfi_input(nchan, 2, nblk, M0, seed=5)
    n = x.shape[2]
    rng = np.random.default_rng(2)
    mask_h = (rng.random((nblk, nchan)) < 0.25).astype(np.uint8)
    sk = dspx.SpectralKurtosis(ctx, nchan, 2, M0)
    sk.set_limits(*rtables[64], 3.0)
    sk.add_resolution(256, 4, *rtables[256])
    np0, nout = sk.blocks_res(n)
    expect, nout_o = oracle.sk_mask_res(x, 64, 1, 256, 4, mask_h)
    assert nout == nout_o
    out = torch.zeros((nchan, 2, n, 2), dtype=torch.float32, device="cuda")
    sk.mask_res(to_dev(mask_h), to_dev(x.view(np.float32)), 2 * n * 2, n * 2, out, 2 * n * 2, n * 2, n)
    got = from_dev(out).view(np.complex64).reshape(nchan, 2, n)
    assert np.array_equal(got[:, :, :nout], expect[:, :, :nout])
    sk.close()
