"""The filterbank engine against the oracle's FilterbankEngineCPU
restatement: a dedispersing filterbank of one complex channel (the
reference's response order for that case) and a multichannel one with a
response table, within 1e-5 relative RMS."""
import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _run(dspx, ctx, x, nsub, fres, pos, neg, resp, batches=None, zdm=False):
    """The engine's output; with zdm, (output, zero-DM output) of one
    dspx_fb_perform_zdm call."""
    import torch
    nchan, npol, ndat = x.shape
    F = dspx.Filterbank(ctx, nchan, npol, nsub, fres, pos, neg)
    F.set_response(resp)
    npart, nout = F.geometry(ndat)
    out = torch.zeros((nchan * nsub, npol, nout, 2), dtype=torch.float32, device="cuda")
    z = torch.zeros_like(out) if zdm else None
    F.perform(to_dev(x.view(np.float32)), npol * ndat * 2, ndat * 2, ndat, out, npol * nout * 2, nout * 2, zdm=z)
    got = from_dev(out).view(np.complex64)[..., 0]
    if batches is not None:
        batches.append(F.batches())
    F.close()
    if zdm:
        return got, from_dev(z).view(np.complex64)[..., 0]
    return got


def test_dedispersing_filterbank_single_channel(ctx, dspx, oracle):
    # -F 64:D on one 16 MHz complex channel at DM 40 (fres from the reference rule)
    cfreq, bw, dm, nsub = 1405.0, 16.0, 40.0, 64
    g = oracle.geometry(cfreq, bw, nsub, dm)
    fres, pos, neg = g["nfft"], g["impulse_pos"], g["impulse_neg"]
    resp = oracle.filterbank_response(cfreq, bw, 1, nsub, dm, fres)
    rng = np.random.default_rng(3)
    ndat = 3 * nsub * fres
    x = (rng.standard_normal((1, 2, ndat)) + 1j * rng.standard_normal((1, 2, ndat))).astype(np.complex64) * 10
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    print("filterbank", nsub, fres, pos, neg, flush=True)
    got = _run(dspx, ctx, x, nsub, fres, pos, neg, resp)
    assert got.shape == expect.shape
    e = rel_rms(got, expect)
    print("rel rms", e, flush=True)
    assert e < 1e-5


def test_filterbank_multichannel_table(ctx, dspx, oracle):
    rng = np.random.default_rng(4)
    nchan, npol, nsub, fres, pos, neg = 3, 2, 16, 128, 10, 12
    nfft = nsub * fres
    resp = (np.exp(1j * rng.uniform(0, 2 * np.pi, (nchan, nfft))) / np.sqrt(nfft * fres)).astype(np.complex64)
    ndat = 5 * nfft + 77
    x = (rng.standard_normal((nchan, npol, ndat)) + 1j * rng.standard_normal((nchan, npol, ndat))).astype(np.complex64)
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    print("filterbank multichannel", flush=True)
    got = _run(dspx, ctx, x, nsub, fres, pos, neg, resp)
    assert rel_rms(got, expect) < 1e-5


def test_filterbank_several_batches(ctx, dspx, oracle, monkeypatch):
    """Windows spread over several library batches, the last one shorter
    (a 1 MB scratch holds 64 windows of 2048 points: 246 windows in batches
    of 62, 62, 62 and 60), against the oracle."""
    rng = np.random.default_rng(6)
    nchan, npol, nsub, fres, pos, neg = 3, 2, 16, 128, 10, 12
    nfft = nsub * fres
    resp = (np.exp(1j * rng.uniform(0, 2 * np.pi, (nchan, nfft))) / np.sqrt(nfft * fres)).astype(np.complex64)
    in_step = nsub * (fres - pos - neg)
    ndat = 40 * in_step + nfft + 5
    x = (rng.standard_normal((nchan, npol, ndat)) + 1j * rng.standard_normal((nchan, npol, ndat))).astype(np.complex64)
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    monkeypatch.setenv("DSPX_FB_SCRATCH_MB", "1")
    monkeypatch.setenv("DSPX_FB_FORM", "library")  # 2048-point windows would otherwise run as one kernel
    print("filterbank several batches", flush=True)
    nb = []
    got = _run(dspx, ctx, x, nsub, fres, pos, neg, resp, nb)
    assert nb == [4]
    assert got.shape == expect.shape and expect.shape[2] == 41 * (fres - pos - neg)
    assert rel_rms(got, expect) < 1e-5


@pytest.mark.parametrize("form", ["fused", "library"])
@pytest.mark.parametrize("nsub,fres", [(16, 64), (8, 128), (8, 256), (32, 128), (16, 256), (64, 64), (32, 32),
                                       (64, 32), (128, 32), (2, 512), (4, 512), (8, 512), (4, 256), (32, 64)])
def test_filterbank_window_lengths(ctx, dspx, oracle, monkeypatch, form, nsub, fres):
    # 1024-4096-point windows: the one-kernel form and the library form
    # (DSPX_FB_FORM=library) against the oracle, 2 channels x 2 pols
    if form == "library":
        monkeypatch.setenv("DSPX_FB_FORM", "library")
    else:
        monkeypatch.delenv("DSPX_FB_FORM", raising=False)
    rng = np.random.default_rng(nsub * 1000 + fres)
    nchan, npol, pos, neg = 2, 2, 9, 14
    nfft = nsub * fres
    resp = (np.exp(1j * rng.uniform(0, 2 * np.pi, (nchan, nfft))) / np.sqrt(nfft * fres)).astype(np.complex64)
    ndat = 4 * nfft + 33
    x = (rng.standard_normal((nchan, npol, ndat)) + 1j * rng.standard_normal((nchan, npol, ndat))).astype(np.complex64)
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    print("filterbank", form, nsub, fres, flush=True)
    nb = []
    got = _run(dspx, ctx, x, nsub, fres, pos, neg, resp, nb)
    assert (nb[0] == 0) == (form == "fused")
    assert got.shape == expect.shape
    e = rel_rms(got, expect)
    print("rel rms", e, flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("nsub,fres,nchan", [(16, 512, 2), (8, 2048, 1), (64, 512, 1), (16, 4096, 2), (8, 32768, 1),
                                             (128, 2048, 1), (512, 512, 1), (2, 1 << 19, 1), (4, 1 << 20, 1)])
def test_filterbank_four_step(ctx, dspx, oracle, monkeypatch, nsub, fres, nchan):
    """Windows of 2^13 to 2^22 points in the four-step form (the rows'
    sub-band transforms in stage B, the sub-band store in stage C), every row
    length (512-4096) and segment lengths from 8 to 2048 points, against the
    oracle and against the library form of the same engine."""
    monkeypatch.delenv("DSPX_FB_FORM", raising=False)
    rng = np.random.default_rng(nsub + fres)
    npol, pos, neg = 2, 21, 17
    nfft = nsub * fres
    resp = (np.exp(1j * rng.uniform(0, 2 * np.pi, (nchan, nfft))) / np.sqrt(nfft * fres)).astype(np.complex64)
    ndat = 3 * nfft + 101
    x = (rng.standard_normal((nchan, npol, ndat)) + 1j * rng.standard_normal((nchan, npol, ndat))).astype(np.complex64)
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    print("four-step filterbank", nsub, fres, nchan, flush=True)
    nb = []
    got = _run(dspx, ctx, x, nsub, fres, pos, neg, resp, nb)
    assert nb == [0]  # not the library form
    assert got.shape == expect.shape
    e = rel_rms(got, expect)
    monkeypatch.setenv("DSPX_FB_FORM", "library")
    lib_out = _run(dspx, ctx, x, nsub, fres, pos, neg, resp)
    e_lib = rel_rms(got, lib_out)
    print("rel rms vs oracle %.3g, vs library form %.3g" % (e, e_lib), flush=True)
    assert e < 1e-5 and e_lib < 1e-5


def test_dedispersing_filterbank_long_windows(ctx, dspx, oracle):
    """-F 16:D on one 16 MHz complex channel at DM 1000: the reference rule's
    sub-band length gives 2^15-bin sub-bands in 2^19-point windows (the
    four-step form), against the oracle with the dedispersion response."""
    cfreq, bw, dm, nsub = 1405.0, 16.0, 1000.0, 16
    g = oracle.geometry(cfreq, bw, nsub, dm)
    fres, pos, neg = g["nfft"], g["impulse_pos"], g["impulse_neg"]
    resp = oracle.filterbank_response(cfreq, bw, 1, nsub, dm, fres)
    rng = np.random.default_rng(5)
    ndat = 3 * nsub * fres
    x = (rng.standard_normal((1, 2, ndat)) + 1j * rng.standard_normal((1, 2, ndat))).astype(np.complex64) * 10
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    print("dedispersing filterbank", nsub, fres, pos, neg, flush=True)
    nb = []
    got = _run(dspx, ctx, x, nsub, fres, pos, neg, resp, nb)
    assert nb == [0]
    assert got.shape == expect.shape
    e = rel_rms(got, expect)
    print("rel rms", e, flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("form,nsub,fres,nchan", [("fused", 16, 128, 2), ("library", 16, 128, 2),
                                                  ("fused", 4, 256, 1), ("four-step", 16, 512, 2),
                                                  ("four-step", 8, 32768, 1), ("library", 8, 2048, 1)])
def test_filterbank_zero_dm_output(ctx, dspx, oracle, monkeypatch, form, nsub, fres, nchan):
    """The zero-DM output (FilterbankEngineCPU.C:266-300: each window's
    sub-bands transformed back without the response, no zero-DM response
    set) in every form: against the oracle's filterbank with a unit response,
    while the dedispersed output of the same call stays that of the plain
    perform."""
    if form == "library":
        monkeypatch.setenv("DSPX_FB_FORM", "library")
    else:
        monkeypatch.delenv("DSPX_FB_FORM", raising=False)
    rng = np.random.default_rng(nsub * 7 + fres)
    npol, pos, neg = 2, 11, 6
    nfft = nsub * fres
    resp = (np.exp(1j * rng.uniform(0, 2 * np.pi, (nchan, nfft))) / np.sqrt(nfft * fres)).astype(np.complex64)
    ndat = 3 * nfft + 57
    x = (rng.standard_normal((nchan, npol, ndat)) + 1j * rng.standard_normal((nchan, npol, ndat))).astype(np.complex64)
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    expect_z = oracle.filterbank(x, nsub, fres, pos, neg, np.ones((nchan, nfft), np.complex64))
    print("filterbank zero-DM", form, nsub, fres, nchan, flush=True)
    nb = []
    got, got_z = _run(dspx, ctx, x, nsub, fres, pos, neg, resp, nb, zdm=True)
    assert (nb[0] > 0) == (form == "library")
    plain = _run(dspx, ctx, x, nsub, fres, pos, neg, resp)
    assert got_z.shape == expect_z.shape == got.shape
    e, ez, ep = rel_rms(got, expect), rel_rms(got_z, expect_z), rel_rms(got, plain)
    print("rel rms: output %.3g, zero-DM %.3g, output vs plain perform %.3g" % (e, ez, ep), flush=True)
    assert e < 1e-5 and ez < 1e-5 and ep < 1e-6
    # not the dedispersed output: the response changes every sample
    assert rel_rms(got_z / np.sqrt(nfft * fres), got) > 0.5
