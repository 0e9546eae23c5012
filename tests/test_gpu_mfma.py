"""Stage B's 4096-point row transforms on the matrix cores (FP16 hi/lo split
operands, FP32 accumulation; stageb_mfma.hpp, environment DSPX_STAGEB=mfma)
against the oracles: the precision configuration (1024 channels of 0.25 MHz
at 800 MHz, DM 750, int16 input with sigma 300) at FFT lengths 2^16..2^22
for channels at the bottom, middle and top of the band, in parity mode
against the FP32 oracle and in truth mode against the FP64 oracle, each
within 1e-5 relative RMS; and a multichannel int8 LoadToFold (Stokes, three
sub-integrations) whose profiles match the oracle chain and whose hits are
bit-exact."""
import numpy as np
import pytest

from helpers import from_dev, oracle_load_to_fold, rel_rms, to_dev

pytestmark = pytest.mark.gpu

CFREQ, BW, NCHAN, DM = 800.0, 256.0, 1024, 750.0


@pytest.mark.parametrize("m,ichan", [(m, c) for m in (16, 18, 20) for c in (0, 511, 1023)] + [(22, 511)])
def test_mfma_precision_configuration(ctx, dspx, oracle, monkeypatch, m, ichan):
    import torch
    monkeypatch.setenv("DSPX_STAGEB", "mfma")
    cw = BW / NCHAN
    fc = oracle.lib().ox_channel_centre(CFREQ, BW, NCHAN, 0, ichan)
    n = 1 << m
    g = oracle.geometry(fc, cw, 1, DM, n)
    pos, neg, step = g["impulse_pos"], g["impulse_neg"], g["step"]
    ndat = n + step
    rng = np.random.default_rng(100 * m + ichan)
    raw = np.clip(np.rint(rng.standard_normal((ndat, 1, 2, 2)) * 300), -32768, 32767).astype(np.int16)
    x = oracle.unpack16(raw.reshape(-1), 1, 2, ndat).view(np.complex64).reshape(1, 2, ndat)
    errs = {}
    for mode, ref in ((dspx.CHIRP_PARITY, 0), (dspx.CHIRP_TRUTH, "f64")):
        resp = oracle.response(fc, cw, 1, DM, n, ref)
        xin = x.astype(np.complex128) if ref == "f64" else x
        expect = oracle.convolve(xin, resp, pos, neg)
        L = dspx.LoadToFold(ctx, 16, 1, 2, ndat, fc, cw, DM, 64, chirp_mode=mode, form=dspx.FORM_TWOLEVEL, geom=g)
        out = torch.zeros((1, 2, L.ndat_out, 2), dtype=torch.float32, device="cuda")
        print("mfma precision m=%d chan=%d mode=%d" % (m, ichan, mode), flush=True)
        L.filterbank(to_dev(raw), out)
        got = from_dev(out).view(np.complex64).reshape(expect.shape)
        errs[mode] = rel_rms(got, expect)
        L.close()
    print("2^%d chan %d: parity vs FP32 oracle %.3g, truth vs FP64 oracle %.3g"
          % (m, ichan, errs[dspx.CHIRP_PARITY], errs[dspx.CHIRP_TRUTH]))
    assert errs[dspx.CHIRP_PARITY] < 1e-5
    assert errs[dspx.CHIRP_TRUTH] < 1e-5


def test_mfma_load_to_fold(ctx, dspx, oracle, monkeypatch):
    import torch
    from importlib import import_module
    monkeypatch.setenv("DSPX_STAGEB", "mfma")
    synth = import_module("dspsr-kanyixia_amd.synth")
    nbit, nchan, npol, ndat, cfreq, bw, dm, nbin, period, nsub = 8, 8, 2, 1 << 19, 1400.0, 25.0, 300.0, 512, 0.0253, 3
    g = oracle.geometry(cfreq, bw, nchan, dm)
    assert g["nfft"] >= (1 << 16)
    raw = synth.make_baseband(nchan, npol, ndat, cfreq, bw, dm, period, 0.03, 2.0, nbit=nbit, seed=5)
    L = dspx.LoadToFold(ctx, nbit, nchan, npol, ndat, cfreq, bw, dm, nbin, state=dspx.STOKES, ndim=4, geom=g)
    rate = bw / nchan * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, period, g["impulse_pos"], nsub=nsub)
    pps = (1.0 / rate) / period
    L.set_fold(pps, chunks)
    fb_o, prof_o, hits_o = oracle_load_to_fold(oracle, raw, nbit, nchan, npol, ndat, cfreq, bw, dm, g, dspx.STOKES, 4,
                                               nbin, pps, chunks, nsub)
    d_raw = to_dev(raw)
    fb = torch.zeros((nchan, npol, L.ndat_out, 2), dtype=torch.float32, device="cuda")
    prof = torch.zeros(prof_o.shape, dtype=torch.float32, device="cuda")
    hits = torch.zeros(hits_o.shape, dtype=torch.int32, device="cuda")
    print("mfma load to fold", g, flush=True)
    L.filterbank(d_raw, fb)
    L.execute(d_raw, prof, hits)
    e_fb = rel_rms(from_dev(fb).view(np.complex64).reshape(fb_o.shape), fb_o)
    e_pr = rel_rms(from_dev(prof), prof_o)
    print("filterbank %.3g profile %.3g" % (e_fb, e_pr), flush=True)
    L.close()
    assert e_fb < 1e-5 and e_pr < 1e-5
    assert np.array_equal(from_dev(hits).view(np.uint32), hits_o)
