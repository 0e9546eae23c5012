"""Inter-channel dispersion delays (dspsr -K) on the GPU: the convolution
with each channel's fractional delay in the on-chip chirp (fused and plain
forms) against the oracle within 1e-5 relative RMS, the SampleDelay engine
bit-exact against the oracle's copy, and a dispersed broadband impulse
aligned to one sample in every channel."""
import math

import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev
from test_interchan import dispersed_impulse

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("nfft", [0, 16384])
def test_interchan_convolution_and_delay(ctx, dspx, oracle, form, nfft):
    import torch
    cfreq, bw, nchan, dm = 1400.0, 32.0, 16, 40.0
    g = oracle.geometry(cfreq, bw, nchan, dm, nfft)
    n, pos, neg = g["nfft"], g["impulse_pos"], g["impulse_neg"]
    ndat = 4 * n
    t0 = ndat // 2
    x1, sd = dispersed_impulse(oracle, cfreq, bw, nchan, dm, ndat, t0)
    rng = np.random.default_rng(5)
    x = np.empty((nchan, 2, ndat), np.complex64)
    x[:, 0] = x1[:, 0]
    x[:, 1] = (rng.standard_normal((nchan, ndat)) + 1j * rng.standard_normal((nchan, ndat))) * 0.1
    resp = oracle.response(cfreq, bw, nchan, dm, n, 0, interchan=True)
    expect = oracle.convolve(x, resp, pos, neg)
    npart = expect.shape[2] // (n - pos - neg)
    nout = expect.shape[2]
    conv = dspx.Convolution(ctx, nchan, 2, n, pos, neg, chirp_mode=dspx.CHIRP_PARITY, form=form,
                            centre_frequency=cfreq, bandwidth=bw, dm=dm, interchan=True)
    out = torch.zeros((nchan, 2, nout, 2), dtype=torch.float32, device="cuda")
    print("interchan conv form", form, "nfft", n, flush=True)
    conv.perform(to_dev(x.view(np.float32)), 2 * ndat * 2, ndat * 2, ndat, out, 2 * nout * 2, nout * 2, npart)
    res = from_dev(out).view(np.complex64).reshape(nchan, 2, nout)
    err = rel_rms(res, expect)
    print("interchan conv rel rms %.3g" % err)
    assert err < 1e-5
    # integer offsets on the device: a copy, so bit-exact against the oracle's copy of the same samples
    nret = nout - sd["total"]
    D = dspx.SampleDelay(ctx, nchan, 2, sd["applied"])
    ret = torch.zeros((nchan, 2, nret, 2), dtype=torch.float32, device="cuda")
    print("retard", nret, flush=True)
    D.retard(out, 2 * nout * 2, nout * 2, 2, nret, ret, 2 * nret * 2, nret * 2)
    got = from_dev(ret).view(np.complex64).reshape(nchan, 2, nret)
    assert np.array_equal(got, oracle.retard(res, sd["applied"], nret))
    # the dispersed impulse is one sample at one index in every channel
    y = np.abs(got[:, 0]) / math.sqrt(n)
    at = t0 - pos - (sd["applied"][0] + sd["delay"][0])
    assert np.all(np.argmax(y, axis=1) == at)
    assert y.max(axis=1).min() > 0.99
    D.close()
    conv.close()


def test_retard_unaligned_rows(ctx, dspx, oracle):
    """Odd offsets and row strides (the single-float path) and ndim 1."""
    import torch
    rng = np.random.default_rng(9)
    nchan, npol, ndat = 5, 2, 1001
    applied = np.array([0, 3, 7, 1, 11], np.int64)
    nret = ndat - int(applied.max())
    x = rng.standard_normal((nchan, npol, ndat)).astype(np.float32)
    D = dspx.SampleDelay(ctx, nchan, npol, applied)
    out = torch.zeros((nchan, npol, nret), dtype=torch.float32, device="cuda")
    print("retard unaligned", flush=True)
    D.retard(to_dev(x), npol * ndat, ndat, 1, nret, out, npol * nret, nret)
    assert np.array_equal(from_dev(out), oracle.retard(x[..., None], applied, nret)[..., 0])
    D.close()
