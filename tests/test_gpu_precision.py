"""The precision configuration (1024 channels of 0.25 MHz at 800 MHz, DM 750,
int16 input with sigma 300, FFT lengths 2^16..2^20): for channels at the
bottom, middle and top of the band, the fused convolution in parity mode
(FP32-rounded chirp phase, as the reference stores it) against the FP32
oracle, and in truth mode (FP64 chirp phase) against the FP64 oracle, each
within 1e-5 relative RMS.  A channel of the 1024-channel band has the chirp
of a one-channel band of the same centre and width, so each runs alone."""
import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu

CFREQ, BW, NCHAN, DM = 800.0, 256.0, 1024, 750.0


@pytest.mark.parametrize("m", [16, 18, 20])
@pytest.mark.parametrize("ichan", [0, 511, 1023])
def test_precision_configuration(ctx, dspx, oracle, m, ichan):
    import torch
    cw = BW / NCHAN
    fc = oracle.lib().ox_channel_centre(CFREQ, BW, NCHAN, 0, ichan)
    n = 1 << m
    g = oracle.geometry(fc, cw, 1, DM, n)
    pos, neg, step = g["impulse_pos"], g["impulse_neg"], g["step"]
    ndat = n + step
    rng = np.random.default_rng(100 * m + ichan)
    raw = np.clip(np.rint(rng.standard_normal((ndat, 1, 2, 2)) * 300), -32768, 32767).astype(np.int16)
    x = oracle.unpack16(raw.reshape(-1), 1, 2, ndat).view(np.complex64).reshape(1, 2, ndat)
    npart = (ndat - pos - neg) // step
    nout = npart * step
    errs = {}
    for mode, ref in ((dspx.CHIRP_PARITY, 0), (dspx.CHIRP_TRUTH, "f64")):
        resp = oracle.response(fc, cw, 1, DM, n, ref)
        xin = x.astype(np.complex128) if ref == "f64" else x
        expect = oracle.convolve(xin, resp, pos, neg)
        conv = dspx.Convolution(ctx, 1, 2, n, pos, neg, chirp_mode=mode, form=dspx.FORM_FUSED,
                                centre_frequency=fc, bandwidth=cw, dm=DM)
        out = torch.zeros((1, 2, nout, 2), dtype=torch.float32, device="cuda")
        print("precision m=%d chan=%d mode=%d" % (m, ichan, mode), flush=True)
        conv.perform(to_dev(x.view(np.float32)), 2 * ndat * 2, ndat * 2, ndat, out, 2 * nout * 2, nout * 2, npart)
        got = from_dev(out).view(np.complex64).reshape(1, 2, nout)
        errs[mode] = rel_rms(got, expect)
        conv.close()
    print("2^%d chan %d: parity vs FP32 oracle %.3g, truth vs FP64 oracle %.3g"
          % (m, ichan, errs[dspx.CHIRP_PARITY], errs[dspx.CHIRP_TRUTH]))
    assert errs[dspx.CHIRP_PARITY] < 1e-5
    assert errs[dspx.CHIRP_TRUTH] < 1e-5
