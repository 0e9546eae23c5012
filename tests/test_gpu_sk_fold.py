"""The SpectralKurtosis-excised LoadToFold (LoadToFold1.C:265-377 with
coherent dedispersion) on the GPU against the oracle chain: SK on the zero-DM
copy of the block (the unpacked input from sample nfilt_pos,
Convolution.C:792-799), the detectors' mask, the dedispersed samples of
flagged blocks zeroed (SpectralKurtosis.C:1465-1566), detection, and the fold
of zeroed data with per-channel hits (Fold.C:911-930).  Masks and hits are
bit-exact, profiles within 1e-5 relative RMS, in every form (plain, three
stages, one pass)."""
import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _tables(M):
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sk_limits.npz"))
    return z["lower%d" % M], z["upper%d" % M], z["sym%d" % M], float(z["pcsd%d" % M])


def _rfi_raw(nchan, npol, ndat, M, t0, seed, tone_frac=0.05, imp_frac=0.01, am_frac=0.01, zero_chan=None):
    """int8 TFP [ndat][nchan][npol][2]: Gaussian noise (sigma 10) with CW
    tones (2-20 sigma) in tone_frac of the channels, broadband impulses
    (5-50 sigma, 1-64 samples) in imp_frac of the M-sample blocks counted
    from input sample t0, 50 Hz-like amplitude modulation in am_frac of the
    channels; zero_chan: a channel of zero samples (which unpack to
    0.5 + 0.5i: constant power)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((ndat, nchan, npol, 2)).astype(np.float32) * 10
    t = np.arange(ndat)
    for c in rng.choice(nchan, max(1, int(tone_frac * nchan)), replace=False):
        a = rng.uniform(2, 20) * 10
        f = rng.uniform(-0.5, 0.5)
        x[:, c, :, 0] += (a * np.cos(2 * np.pi * f * t))[:, None]
        x[:, c, :, 1] += (a * np.sin(2 * np.pi * f * t))[:, None]
    nblk = (ndat - t0) // M
    for b in rng.choice(nblk, max(1, int(imp_frac * nblk)), replace=False):
        n = int(rng.integers(1, min(64, M) + 1))
        s = t0 + b * M + int(rng.integers(0, M - n + 1))
        x[s:s + n] *= rng.uniform(5, 50)
    for c in rng.choice(nchan, max(1, int(am_frac * nchan)), replace=False):
        x[:, c] *= (1 + 0.9 * np.sin(2 * np.pi * t / 4096.0))[:, None, None]
    if zero_chan is not None:
        x[:, zero_chan] = 0
    return np.clip(np.rint(x), -127, 127).astype(np.int8)


def _oracle_chain(oracle, raw, nchan, npol, ndat, g, cfreq, bw, dm, state, ndim, nbin, pps, chunks, nsub, M,
                  mask, scale=1.0):
    """Convolution -> SpectralKurtosis::mask -> Detection -> Fold (zeroed
    data, hits_nchan = nchan) on the oracle."""
    x = oracle.unpack8(raw, nchan, npol, ndat, scale=scale)
    x = np.ascontiguousarray(x).view(np.complex64).reshape(nchan, npol, ndat)
    resp = oracle.response(cfreq, bw, nchan, dm, g["nfft"], 0)
    fb = oracle.convolve(x, resp, g["impulse_pos"], g["impulse_neg"])
    cleaned = oracle.sk_mask(fb, M, mask)
    det = oracle.detect(cleaned, state, ndim)
    nout = fb.shape[2]
    npol_o = det.shape[1]
    nd = det.shape[2] // nout
    prof = np.zeros((nsub, nchan, npol_o, nbin * nd), np.float32)
    hits = np.zeros((nsub, nchan, nbin), np.uint32)
    nfold = 0
    for (i0, n, phi, isub) in chunks:
        _, _, k = oracle.fold(det, nd, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof[isub],
                              hits=hits[isub], zeroed=True, hits_nchan=nchan)
        nfold += k
    return x, prof, hits, nfold


def _run(ctx, dspx, L, raw, mask_dev, M, prof_shape, hits_shape):
    import torch
    prof = torch.zeros(prof_shape, dtype=torch.float32, device="cuda")
    hits = torch.zeros(hits_shape, dtype=torch.int32, device="cuda")
    L.execute_masked(to_dev(raw), mask_dev, M, prof, hits)
    return from_dev(prof), from_dev(hits).view(np.uint32)


def test_sk_fold_rfi_band(ctx, dspx, oracle):
    """512 channels x 2 pols of 8-bit samples with CW tones, impulses and
    amplitude modulation, SK at M = 1024 with 3-sigma thresholds (all three
    detectors, omit_outliers: the blocks detect_skfb zaps leave the other
    two detectors' statistics, so that channel totals fall between M and
    32768 and detect_tscr evaluates SKLimits for them), then dedispersion,
    Stokes detection and a fold of 256 bins into two sub-integrations.  The
    GPU's own SK (from the raw block) makes the mask the fold applies."""
    import torch
    from importlib import import_module
    S = import_module("dspsr-kanyixia_amd.sklimits")
    nchan, npol, M, nbin, nsub = 512, 2, 1024, 256, 2
    cfreq, bw, dm, period = 1400.0, 400.0, 50.0, 0.0231
    g = oracle.geometry(cfreq, bw, nchan, dm)
    ndat = (1 << 16) + g["noverlap"]
    pos = g["impulse_pos"]
    raw = _rfi_raw(nchan, npol, ndat, M, pos, seed=17, zero_chan=9)
    L = dspx.LoadToFold(ctx, 8, nchan, npol, ndat, cfreq, bw, dm, nbin, state=dspx.STOKES, ndim=4, geom=g)
    nblk = L.ndat_out // M
    assert nblk >= 32
    lower, upper, sym, pcsd = _tables(M)
    # the oracle: SK of the zero-DM copy (input samples from nfilt_pos)
    x = oracle.unpack8(raw, nchan, npol, ndat)
    x = np.ascontiguousarray(x).view(np.complex64).reshape(nchan, npol, ndat)
    sums_o, counts_o = oracle.sk_compute(np.ascontiguousarray(x[:, :, pos:pos + nblk * M]), M)
    _, _, c_after = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=3, omit_outliers=True)
    totals = sorted({int(t) for t in c_after.astype(np.uint64).sum(axis=0) if M < t < 32768})
    assert totals
    lim = {}
    for t in totals:
        L_t = S.SKLimits(t, 3.0, 1, 2)
        lim[t] = (L_t.lower, L_t.upper)
    mask_o, _, _ = oracle.sk_detect(sums_o, counts_o, M, lower, upper, sym, pcsd, flags=7, omit_outliers=True,
                                    tscr_limits=lim)
    assert 0.01 < mask_o.mean() < 0.5 and mask_o.all(axis=0).sum() >= 16
    # the GPU: SK from the raw block at t0 = nfilt_pos, then the masked chain
    sk = dspx.SpectralKurtosis(ctx, nchan, npol, M, omit_outliers=True)
    sk.set_limits(lower, upper, sym, pcsd, std_devs=3.0)
    d_raw = to_dev(raw)
    sums = torch.zeros((nblk, nchan, npol, 2), dtype=torch.float32, device="cuda")
    counts = torch.zeros((nblk, nchan), dtype=torch.int32, device="cuda")
    mask = torch.zeros((nblk, nchan), dtype=torch.uint8, device="cuda")
    print("sk fold rfi band: sk", g, flush=True)
    sk.compute_raw(d_raw, 8, pos, nblk, sums, counts)
    sk.detect(sums, counts, nblk, 7, mask)
    got_mask = from_dev(mask)
    assert np.array_equal(got_mask, mask_o)
    rate = bw / nchan * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, period, pos, nsub=nsub)
    pps = (1.0 / rate) / period
    L.set_fold(pps, chunks)
    _, prof_o, hits_o, nfold = _oracle_chain(oracle, raw, nchan, npol, ndat, g, cfreq, bw, dm, dspx.STOKES, 4, nbin,
                                            pps, chunks, nsub, M, mask_o)
    print("sk fold rfi band: fold", flush=True)
    prof, hits = _run(ctx, dspx, L, raw, mask, M, prof_o.shape, hits_o.shape)
    e = rel_rms(prof, prof_o)
    print("profile %.3g, folded %d of %d" % (e, nfold, nchan * L.ndat_out), flush=True)
    L.close()
    sk.close()
    assert np.array_equal(hits, hits_o)
    # the channel of constant power is zapped throughout (SK = 0)
    assert hits.sum() == nfold and hits[:, 9].sum() == 0
    assert e < 1e-5


FORM_CASES = [
    # form, nfft (0: the smear rule), state, ndim, nbin, nsub, M (not a power of two), blocks dropped at the end
    ("ONEPASS", 1024, 3, 4, 512, 3, 100, 3),
    ("ONEPASS", 256, 0, 1, 128, 2, 37, 0),
    ("TWOLEVEL", 1 << 14, 3, 4, 256, 3, 1000, 2),
    ("TWOLEVEL", 1 << 13, 1, 1, 1024, 2, 64, 1),
    ("PLAIN", 1 << 13, 3, 4, 256, 3, 1000, 2),
    ("FUSED", 0, 2, 1, 64, 2, 512, 1),
]


@pytest.mark.parametrize("form,nfft,state,ndim,nbin,nsub,M,drop", FORM_CASES)
def test_sk_fold_forms(ctx, dspx, oracle, form, nfft, state, ndim, nbin, nsub, M, drop):
    """Random masks over blocks of M output samples (M not a power of two;
    the last blocks left out, as past the estimator's blocks), sub-integration
    edges inside windows: every form's fold against the oracle chain."""
    nchan, npol = 12, 2
    cfreq, bw, dm, period = 1500.0, 24.0, 8.0, 0.0047
    g = dspx.geometry(cfreq, bw, nchan, dm, nfft=nfft)
    ndat = 40 * g["step"] + g["noverlap"] + 11
    raw = _rfi_raw(nchan, npol, ndat, M, g["impulse_pos"], seed=M + nbin)
    L = dspx.LoadToFold(ctx, 8, nchan, npol, ndat, cfreq, bw, dm, nbin, state=state, ndim=ndim,
                        form=getattr(dspx, "FORM_" + form), geom=g)
    nblk = L.ndat_out // M - drop
    rng = np.random.default_rng(nbin)
    mask = (rng.random((nblk, nchan)) < 0.2).astype(np.uint8)
    mask[:, 7] = 1  # a whole channel zapped
    rate = bw / nchan * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, period, g["impulse_pos"], nsub=nsub)
    pps = (1.0 / rate) / period
    L.set_fold(pps, chunks)
    _, prof_o, hits_o, nfold = _oracle_chain(oracle, raw, nchan, npol, ndat, g, cfreq, bw, dm, state, ndim, nbin,
                                            pps, chunks, nsub, M, mask)
    print("sk fold form %s nfft %d state %d M %d nblk %d" % (form, g["nfft"], state, M, nblk), flush=True)
    prof, hits = _run(ctx, dspx, L, raw, to_dev(mask), M, prof_o.shape, hits_o.shape)
    e = rel_rms(prof, prof_o)
    print("profile %.3g" % e, flush=True)
    L.close()
    assert np.array_equal(hits, hits_o)
    assert hits[:, 7].sum() == 0 and hits.sum() == nfold
    assert e < 1e-5


@pytest.mark.parametrize("form", ["ONEPASS", "TWOLEVEL", "PLAIN"])
def test_sk_fold_zero_values_not_counted(ctx, dspx, oracle, form):
    """Unmasked samples whose detected value is exactly zero (a block
    unpacked with scale 0) add no hits (Fold.C:911-930), in every form."""
    nchan, npol, M, nbin = 4, 2, 64, 128
    g = dspx.geometry(1500.0, 24.0, nchan, 8.0, nfft=1024 if form == "ONEPASS" else 1 << 13)
    ndat = 6 * g["step"] + g["noverlap"]
    raw = _rfi_raw(nchan, npol, ndat, M, g["impulse_pos"], seed=5)
    L = dspx.LoadToFold(ctx, 8, nchan, npol, ndat, 1500.0, 24.0, 8.0, nbin, state=dspx.STOKES, ndim=4,
                        form=getattr(dspx, "FORM_" + form), scale=0.0, geom=g)
    mask = np.zeros((L.ndat_out // M, nchan), np.uint8)
    mask[::3, 1] = 1
    rate = 24.0 / nchan * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, 0.0031, g["impulse_pos"], nsub=2)
    pps = (1.0 / rate) / 0.0031
    L.set_fold(pps, chunks)
    _, prof_o, hits_o, nfold = _oracle_chain(oracle, raw, nchan, npol, ndat, g, 1500.0, 24.0, 8.0, dspx.STOKES, 4,
                                            nbin, pps, chunks, 2, M, mask, scale=0.0)
    assert nfold == 0 and not hits_o.any() and not prof_o.any()
    print("sk fold zero values", form, flush=True)
    prof, hits = _run(ctx, dspx, L, raw, to_dev(mask), M, prof_o.shape, hits_o.shape)
    L.close()
    assert not hits.any() and not prof.any()


def test_sk_fold_rejects_bad_mask_shape(ctx, dspx, oracle):
    import torch
    g = dspx.geometry(1500.0, 24.0, 4, 8.0, nfft=1024)
    ndat = 10 * g["step"] + g["noverlap"]
    L = dspx.LoadToFold(ctx, 8, 4, 2, ndat, 1500.0, 24.0, 8.0, 64, geom=g)
    L.set_fold(1e-4, [(0, L.ndat_out, 0.0, 0)])
    raw = torch.zeros((ndat, 4, 2, 2), dtype=torch.int8, device="cuda")
    prof = torch.zeros((1, 4, 1, 64), dtype=torch.float32, device="cuda")
    too_many = torch.zeros((L.ndat_out // 16 + 1, 4), dtype=torch.uint8, device="cuda")
    with pytest.raises(dspx.DspxError) as e:
        L.execute_masked(raw, too_many, 16, prof)
    assert e.value.code == dspx.E_PARAM
    with pytest.raises(dspx.DspxError):
        L.execute_masked(raw, too_many[:4], 1, prof)
    L.close()
