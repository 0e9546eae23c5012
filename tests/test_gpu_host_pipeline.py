"""The C++ host layer (dsp::HIP engines behind the dsp Transformations)
driven by dspx_pipeline, against the oracle chain."""
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_load_to_fold, rel_rms
from test_gpu_ltf import CASES, _make

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "dspsr-kanyixia_amd", "dspx_pipeline")
STATES = {0: "Intensity", 1: "PPQQ", 2: "Coherence", 3: "Stokes"}


@pytest.mark.parametrize("mode", ["chain", "fused", "plain"])
@pytest.mark.parametrize("case", ["single_chan_intensity", "multichan_stokes_subints"])
def test_host_pipeline(dspx, oracle, tmp_path, case, mode):
    assert os.path.exists(EXE), "dspx_pipeline not built"
    nbit, nchan, npol, ndat, cfreq, bw, dm, state, ndim, nbin, period, nsub = CASES[case]
    raw = _make(dspx, case)
    rawf = tmp_path / "raw.bin"
    raw.tofile(rawf)
    outf = tmp_path / "prof.bin"
    cmd = [EXE, "--raw", str(rawf), "--nbit", str(nbit), "--nchan", str(nchan), "--npol", str(npol),
           "--ndat", str(ndat), "--cfreq", repr(cfreq), "--bw", repr(bw), "--dm", repr(dm), "--period",
           repr(period), "--nbin", str(nbin), "--nsub", str(nsub), "--state", STATES[state], "--ndim",
           str(ndim), "--mode", mode, "--out", str(outf)]
    print("host pipeline", case, mode, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    geom = oracle.geometry(cfreq, bw, nchan, dm)
    step = geom["step"]
    nout = ((ndat - geom["noverlap"]) // step) * step
    rate = abs(bw) / nchan * 1e6
    chunks = dspx.fold_chunks(nout, rate, period, geom["impulse_pos"], nsub=nsub)
    pps = (1.0 / rate) / period
    _, prof_o, hits_o = oracle_load_to_fold(oracle, raw, nbit, nchan, npol, ndat, cfreq, bw, dm, geom, state,
                                            ndim, nbin, pps, chunks, nsub)
    data = np.fromfile(outf, np.uint8)
    nprof = prof_o.size
    prof = data[:nprof * 4].view(np.float32).reshape(prof_o.shape)
    hits = data[nprof * 4:].view(np.uint32).reshape(hits_o.shape)
    e = rel_rms(prof, prof_o)
    print("profile rel rms %.3g" % e)
    assert e < 1e-5
    assert np.array_equal(hits, hits_o)


@pytest.mark.parametrize("with_response", [False, True])
def test_host_filterbank(dspx, oracle, tmp_path, with_response):
    """Unpacker -> Filterbank[HIP::FilterbankEngine] through the C++ host
    layer, against the oracle's FilterbankEngineCPU restatement on the same
    unpacked samples: 2 input channels x 16 sub-bands of 128 bins, with no
    response (the normaliser alone) and with a response table; with the
    response, the zero-DM output too (the sub-bands without the response or
    the normaliser)."""
    assert os.path.exists(EXE), "dspx_pipeline not built"
    rng = np.random.default_rng(17)
    nchan, npol, nsub, fres = 2, 2, 16, 128
    pos, neg = (10, 12) if with_response else (0, 0)
    nfft = nsub * fres
    ndat = 6 * nfft + 300
    raw = rng.integers(-60, 61, (ndat, nchan, npol, 2)).astype(np.int8)
    rawf = tmp_path / "raw.bin"
    raw.tofile(rawf)
    outf = tmp_path / "fb.bin"
    cmd = [EXE, "--raw", str(rawf), "--nbit", "8", "--nchan", str(nchan), "--npol", str(npol), "--ndat", str(ndat),
           "--cfreq", "1400.0", "--bw", "64.0", "--dm", "0.0", "--mode", "filterbank", "--fb_nchan",
           str(nchan * nsub), "--fres", str(fres), "--out", str(outf)]
    norm = np.float32(1.0 / np.sqrt(float(nfft) * fres))
    if with_response:
        table = np.exp(1j * rng.uniform(0, 2 * np.pi, (nchan * nsub, fres))).astype(np.complex64)
        respf = tmp_path / "resp.bin"
        table.tofile(respf)
        zdmf = tmp_path / "zdm.bin"
        cmd += ["--response", str(respf), "--nfilt_pos", str(pos), "--nfilt_neg", str(neg), "--zdm", str(zdmf)]
        resp = (table.view(np.float32) * norm).view(np.complex64).reshape(nchan, nfft)
    else:
        resp = np.full((nchan, nfft), norm, np.complex64)
    print("host filterbank response=%s" % with_response, flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    x = oracle.unpack8(raw.reshape(-1), nchan, npol, ndat).view(np.complex64)[..., 0]
    expect = oracle.filterbank(x, nsub, fres, pos, neg, resp)
    got = np.fromfile(outf, np.float32).view(np.complex64).reshape(expect.shape)
    e = rel_rms(got, expect)
    print("filterbank rel rms %.3g" % e)
    assert e < 1e-5
    if with_response:
        expect_z = oracle.filterbank(x, nsub, fres, pos, neg, np.ones((nchan, nfft), np.complex64))
        got_z = np.fromfile(zdmf, np.float32).view(np.complex64).reshape(expect_z.shape)
        ez = rel_rms(got_z, expect_z)
        print("zero-DM rel rms %.3g" % ez)
        assert ez < 1e-5


def test_host_pipeline_interchan(dspx, oracle, tmp_path):
    """The chain with inter-channel delays removed (dspsr -K):
    Convolution[fractional-delay chirp] -> SampleDelay -> Detection -> Fold
    through the C++ host layer, against the oracle's chain."""
    assert os.path.exists(EXE), "dspx_pipeline not built"
    case = "multichan_stokes_subints"
    nbit, nchan, npol, ndat, cfreq, bw, dm, state, ndim, nbin, period, nsub = CASES[case]
    raw = _make(dspx, case)
    rawf = tmp_path / "raw.bin"
    raw.tofile(rawf)
    outf = tmp_path / "prof.bin"
    cmd = [EXE, "--raw", str(rawf), "--nbit", str(nbit), "--nchan", str(nchan), "--npol", str(npol),
           "--ndat", str(ndat), "--cfreq", repr(cfreq), "--bw", repr(bw), "--dm", repr(dm), "--period",
           repr(period), "--nbin", str(nbin), "--nsub", str(nsub), "--state", STATES[state], "--ndim",
           str(ndim), "--mode", "chain", "--interchan", "1", "--out", str(outf)]
    print("host pipeline interchan", flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    geom = oracle.geometry(cfreq, bw, nchan, dm)
    x = oracle.unpack8(raw, nchan, npol, ndat).view(np.complex64).reshape(nchan, npol, ndat)
    resp = oracle.response(cfreq, bw, nchan, dm, geom["nfft"], 0, interchan=True)
    fb = oracle.convolve(x, resp, geom["impulse_pos"], geom["impulse_neg"])
    sd = oracle.sample_delays(cfreq, bw, nchan, dm)
    nret = fb.shape[2] - sd["total"]
    assert nret > 0
    det = oracle.detect(oracle.retard(fb, sd["applied"], nret), state, ndim)
    rate = abs(bw) / nchan * 1e6
    pps = (1.0 / rate) / period
    chunks = dspx.fold_chunks(nret, rate, period, geom["impulse_pos"] + sd["zero_delay"], nsub=nsub)
    nd = det.shape[2] // nret
    prof_o = np.zeros((nsub, nchan, det.shape[1], nbin * nd), np.float32)
    hits_o = np.zeros((nsub, 1, nbin), np.uint32)
    for (i0, n, phi, isub) in chunks:
        oracle.fold(det, nd, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof_o[isub], hits=hits_o[isub])
    data = np.fromfile(outf, np.uint8)
    nprof = prof_o.size
    prof = data[:nprof * 4].view(np.float32).reshape(prof_o.shape)
    hits = data[nprof * 4:].view(np.uint32).reshape(nsub, nbin)
    e = rel_rms(prof, prof_o)
    print("interchan profile rel rms %.3g" % e)
    assert e < 1e-5
    assert np.array_equal(hits, hits_o[:, 0])


@pytest.mark.parametrize("mode", ["chain", "fused"])
def test_host_pipeline_dada(dspx, oracle, tmp_path, mode):
    """The same block read from a DADA file (shape, band and sample rate from
    its header) gives the raw run's profiles and hits (bit-identical in the
    fused form), which the oracle chain checks; the .npz archive holds them."""
    from test_dada import write_dada
    assert os.path.exists(EXE), "dspx_pipeline not built"
    case = "multichan_stokes_subints"
    nbit, nchan, npol, ndat, cfreq, bw, dm, state, ndim, nbin, period, nsub = CASES[case]
    raw = _make(dspx, case)
    tsamp = nchan / abs(bw)  # microseconds per complex sample of a channel
    assert 1e6 / tsamp == abs(bw) / nchan * 1e6
    dada = tmp_path / "block.dada"
    write_dada(dada, raw.reshape(ndat, nchan, npol, 2), cfreq, bw, tsamp, nbit)
    rawf = tmp_path / "raw.bin"
    raw.tofile(rawf)
    common = ["--dm", repr(dm), "--period", repr(period), "--nbin", str(nbin), "--nsub", str(nsub), "--state",
              STATES[state], "--ndim", str(ndim), "--mode", mode]
    out_dada, out_raw = tmp_path / "dada.bin", tmp_path / "raw_out.bin"
    print("host pipeline dada", mode, flush=True)
    npz = tmp_path / "fold.npz"
    r = subprocess.run([EXE, "--dada", str(dada), "--out", str(out_dada), "--npz", str(npz)] + common,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    r = subprocess.run([EXE, "--raw", str(rawf), "--nbit", str(nbit), "--nchan", str(nchan), "--npol", str(npol),
                        "--ndat", str(ndat), "--cfreq", repr(cfreq), "--bw", repr(bw), "--out", str(out_raw)] + common,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0
    a, b = np.fromfile(out_dada, np.uint8), np.fromfile(out_raw, np.uint8)
    if mode == "fused":
        assert np.array_equal(a, b)  # fixed-order reductions
    else:
        # the Fold engine adds with atomics (as the reference's does): the
        # order of the float additions, and so the last bits, vary run to run
        nb = nsub * nbin * 4
        assert np.array_equal(a[-nb:], b[-nb:])
        assert rel_rms(a[:-nb].view(np.float32), b[:-nb].view(np.float32)) < 1e-6
    # the archive holds the same profiles and hits, with the fold's parameters
    z = np.load(npz, allow_pickle=False)
    data = np.fromfile(out_dada, np.uint8)
    nprof = z["profiles"].size
    assert z["profiles"].shape == (nsub, nchan, 4 // ndim, nbin * ndim)
    assert np.array_equal(z["profiles"].ravel(), data[:nprof * 4].view(np.float32))
    assert np.array_equal(z["hits"].ravel(), data[nprof * 4:].view(np.uint32))
    assert float(z["folding_period"]) == period and float(z["dispersion_measure"]) == dm
