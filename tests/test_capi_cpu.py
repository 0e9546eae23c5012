"""The C ABI without a device: the library loads, exports every function
include/dspx.h declares, and its host-side algorithms (geometry, bin plan)
match the oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dspx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dspx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(dspx):
    L = dspx.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # the Python binding covers the whole ABI
    assert set(names) == set(dspx.exported_symbols())
    assert dspx.lib().dspx_version().decode().startswith("dspx")


@pytest.mark.parametrize("cfreq,bw,nchan,dm,nfft", [
    (1405.0, 16.0, 1, 56.77, 0), (1400.0, 400.0, 128, 300.0, 0), (800.0, 256.0, 1024, 750.0, 0),
    (300.0, 75.0, 64, 1500.0, 0), (1500.0, 375.0, 512, 5.0, 0), (1400.0, -400.0, 128, 300.0, 0),
    (800.0, 256.0, 1024, 750.0, 1 << 16), (430.0, 20.0, 32, 71.0, 0)])
def test_geometry_matches_oracle(dspx, oracle, cfreq, bw, nchan, dm, nfft):
    assert dspx.geometry(cfreq, bw, nchan, dm, nfft) == oracle.geometry(cfreq, bw, nchan, dm, nfft)


def test_geometry_errors_like_the_reference(dspx):
    # every channel's smear beyond the limit: PlasmaResponse::prepare throws
    with pytest.raises(dspx.DspxError) as e:
        dspx.geometry(100.0, 50.0, 1, 5000.0)
    assert e.value.code == dspx.E_STATE
    # user-set nfft smaller than the overlap: Convolution::prepare throws
    with pytest.raises(dspx.DspxError) as e:
        dspx.geometry(1405.0, 16.0, 1, 56.77, nfft=1024)
    assert e.value.code == dspx.E_STATE


@pytest.mark.parametrize("seed", range(6))
def test_binplan_host_matches_serial_recurrence(dspx, oracle, seed):
    rng = np.random.default_rng(100 + seed)
    n = 600_000
    pps = float(10 ** rng.uniform(-7, -2.5))
    phi = float(rng.uniform(-4, 4))
    nbin = int(rng.choice([128, 1000, 1024, 2048, 333]))
    _, _, _, plan = oracle.fold(np.zeros((1, 1, n), np.float32), 1, phi, pps, nbin, want_binplan=True)
    assert np.array_equal(dspx.binplan_host(phi, pps, nbin, n), plan)


def test_binplan_host_single_channel_phase_rate(dspx, oracle):
    """The single-channel workload's rate: 62.5 ns samples folded at 89.33 ms."""
    pps = 62.5e-9 / 0.08933
    for phi in (0.0, 0.25, 0.9999999999):
        _, _, _, plan = oracle.fold(np.zeros((1, 1, 3_000_000), np.float32), 1, phi, pps, 1024,
                                    want_binplan=True)
        assert np.array_equal(dspx.binplan_host(phi, pps, 1024, 3_000_000), plan)


def test_binplan_rejects_bad_rate(dspx):
    with pytest.raises(dspx.DspxError):
        dspx.binplan_host(0.0, 0.7, 64, 10)
