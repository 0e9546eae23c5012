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
