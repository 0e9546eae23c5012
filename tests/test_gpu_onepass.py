"""The one-pass LoadToFold form (N = 2^7 .. 2^13, onepass.hpp) against the
oracle chain (unpack -> Convolution -> Detection -> Fold): the filterbank
output within 1e-5 relative RMS, profiles within 1e-5, hits bit-exact, for
every length, both sample widths, the four detection states and folds of
several sub-integrations whose boundaries fall inside windows and inside
the workgroups' window ranges."""
import numpy as np
import pytest

from helpers import from_dev, oracle_load_to_fold, rel_rms, to_dev

pytestmark = pytest.mark.gpu

CFREQ, BW, NCHAN, DM = 1500.0, 8.0, 4, 2.0  # 2 MHz channels: an overlap of 24 samples


def _run(ctx, dspx, oracle, synth, m, nbit, state, ndim, nbin, nsub, period, form, npart=40, seed=0):
    """One block through the plan in both outputs; returns the stage names of
    the fold execute."""
    import torch
    n = 1 << m
    g = dspx.geometry(CFREQ, BW, NCHAN, DM, nfft=n)
    assert g["nfft"] == n
    ndat = npart * g["step"] + g["noverlap"] + 7  # a ragged tail
    raw = synth.make_baseband(NCHAN, 2, ndat, CFREQ, BW, DM, period, 0.05, 2.0, sigma=10.0 if nbit == 8 else 300.0,
                              nbit=nbit, seed=seed + m)
    L = dspx.LoadToFold(ctx, nbit, NCHAN, 2, ndat, CFREQ, BW, DM, nbin, state=state, ndim=ndim, form=form, geom=g)
    assert L.npart == npart
    rate = BW / NCHAN * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, period, g["impulse_pos"], nsub=nsub)
    pps = (1.0 / rate) / period
    L.set_fold(pps, chunks)
    fb_o, prof_o, hits_o = oracle_load_to_fold(oracle, raw, nbit, NCHAN, 2, ndat, CFREQ, BW, DM, g, state, ndim,
                                               nbin, pps, chunks, nsub)
    d_raw = to_dev(raw)
    fb = torch.zeros((NCHAN, 2, L.ndat_out, 2), dtype=torch.float32, device="cuda")
    prof = torch.zeros(prof_o.shape, dtype=torch.float32, device="cuda")
    hits = torch.zeros(hits_o.shape, dtype=torch.int32, device="cuda")
    print("onepass 2^%d nbit %d state %d ndim %d nbin %d nsub %d form %d" % (m, nbit, state, ndim, nbin, nsub, form),
          flush=True)
    L.filterbank(d_raw, fb)
    L.set_timing(True)
    L.execute(d_raw, prof, hits)
    names = [n for n, _ in L.stage_times()]
    e_fb = rel_rms(from_dev(fb).view(np.complex64).reshape(fb_o.shape), fb_o)
    e_pr = rel_rms(from_dev(prof), prof_o)
    got_hits = from_dev(hits).view(np.uint32)
    print("filterbank %.3g profile %.3g" % (e_fb, e_pr), flush=True)
    L.close()
    assert e_fb < 1e-5
    assert np.array_equal(got_hits, hits_o)
    assert e_pr < 1e-5
    return names


@pytest.mark.parametrize("m", list(range(7, 14)))
def test_onepass_every_length(ctx, dspx, oracle, m):
    synth = pytest.importorskip("dspsr-kanyixia_amd.synth")
    _run(ctx, dspx, oracle, synth, m, 8, dspx.STOKES, 4, 1024, 3, 0.0037, dspx.FORM_ONEPASS)


@pytest.mark.parametrize("nbit,state,ndim", [(16, 0, 1), (8, 1, 1), (16, 2, 2), (8, 2, 1)])
def test_onepass_states(ctx, dspx, oracle, nbit, state, ndim):
    synth = pytest.importorskip("dspsr-kanyixia_amd.synth")
    _run(ctx, dspx, oracle, synth, 9, nbit, state, ndim, 512, 2, 0.011, dspx.FORM_ONEPASS, seed=state)


def test_onepass_chosen_by_fused_form(ctx, dspx, oracle):
    """FORM_FUSED picks the one-pass kernel at 2^12 (and the result is the
    same as FORM_TWOLEVEL's, the three stages, at that length)."""
    synth = pytest.importorskip("dspsr-kanyixia_amd.synth")
    names = _run(ctx, dspx, oracle, synth, 12, 8, dspx.STOKES, 4, 256, 2, 0.005, dspx.FORM_FUSED, npart=12)
    assert any(n.startswith("onepass") for n in names), names
    names = _run(ctx, dspx, oracle, synth, 12, 8, dspx.STOKES, 4, 256, 2, 0.005, dspx.FORM_TWOLEVEL, npart=12)
    assert "stageB_rowfft_chirp" in names, names


def test_onepass_many_windows_per_workgroup(ctx, dspx, oracle):
    """Enough windows that each workgroup loops over several in-flight
    batches (2^8 points: 16 windows in flight, about 2900 windows per
    channel), with eight sub-integrations."""
    synth = pytest.importorskip("dspsr-kanyixia_amd.synth")
    _run(ctx, dspx, oracle, synth, 8, 8, dspx.STOKES, 4, 128, 8, 0.0021, dspx.FORM_ONEPASS, npart=2900)
