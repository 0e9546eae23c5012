"""The benchmark workloads' own shapes against the oracle chain, at reduced
sample counts: the 128-channel band (3.125 MHz channels at 1400 MHz, DM 300:
nfft 2^18 with an overlap of 15418, full Stokes folded into 1024 bins x 8
sub-integrations: the stage-C kernel the multichannel benchmark times), and
a slice of the uneven batch's own draw (seed 3) with its 512-channel,
16-bit, shortest-transform and longest-transform (2^21) beams.  The GPU runs
every channel; the oracle runs a spread of channels (each channel's chain is
independent: its response is the band's response row, bit for bit), whose
profiles must agree within 1e-5 relative RMS, and the hits (the same for
every channel) bit-exactly."""
import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu


def _noise(ndat, nchan, npol, nbit, seed):
    """Uniform integer noise, TFP [ndat][nchan][npol][2]."""
    rng = np.random.default_rng(seed)
    if nbit == 8:
        return rng.integers(-40, 41, (ndat, nchan, npol, 2), dtype=np.int8)
    return rng.integers(-1000, 1001, (ndat, nchan, npol, 2), dtype=np.int16)


def oracle_channels(oracle, raw, nbit, chans, nchan, npol, ndat, cfreq, bw, dm, nfft, pos, neg, state, ndim, nbin,
                    pps, chunks, nsub):
    """The oracle chain (unpack, convolution with the channel's response,
    detection, fold) for the listed channels of a band: profiles
    [nsub][len(chans)][npol_out][nbin*ndim] and hits [nsub][nbin]."""
    sub = np.ascontiguousarray(raw[:, chans])
    k = len(chans)
    x = oracle.unpack8(sub, k, npol, ndat) if nbit == 8 else oracle.unpack16(sub, k, npol, ndat)
    x = np.ascontiguousarray(x).view(np.complex64).reshape(k, npol, ndat)
    cw = bw / nchan
    resp = np.concatenate([oracle.response(oracle.lib().ox_channel_centre(cfreq, bw, nchan, 0, c), cw, 1, dm, nfft, 0)
                           for c in chans])
    fb = oracle.convolve(x, resp, pos, neg)
    det = oracle.detect(fb, state, ndim)
    nout = fb.shape[2]
    npol_o = det.shape[1]
    nd = det.shape[2] // nout
    prof = np.zeros((nsub, k, npol_o, nbin * nd), np.float32)
    hits = np.zeros((nsub, 1, nbin), np.uint32)
    for (i0, n, phi, isub) in chunks:
        oracle.fold(det, nd, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof[isub], hits=hits[isub])
    return prof, hits[:, 0]


def test_config_multichannel_band(ctx, dspx, oracle):
    """128 channels x 2 pols of 8-bit samples, 400 MHz at 1400 MHz, DM 300
    (nfft 2^18, overlap 15418), Stokes ndim 4, 1024 bins x 8
    sub-integrations, two windows per channel."""
    import torch
    nchan, npol, cfreq, bw, dm, nbin, nsub, period = 128, 2, 1400.0, 400.0, 300.0, 1024, 8, 0.253
    g = oracle.geometry(cfreq, bw, nchan, dm)
    assert g["nfft"] == 1 << 18 and g["noverlap"] == 15418
    ndat = 2 * g["step"] + g["noverlap"] + 5
    raw = _noise(ndat, nchan, npol, 8, seed=2)
    L = dspx.LoadToFold(ctx, 8, nchan, npol, ndat, cfreq, bw, dm, nbin, state=dspx.STOKES, ndim=4, geom=g)
    assert L.npart == 2
    rate = bw / nchan * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, period, g["impulse_pos"], nsub=nsub)
    pps = (1.0 / rate) / period
    L.set_fold(pps, chunks)
    prof = torch.zeros((nsub, nchan, 1, 4 * nbin), dtype=torch.float32, device="cuda")
    hits = torch.zeros((nsub, nbin), dtype=torch.int32, device="cuda")
    print("config multichannel band", g, flush=True)
    L.set_timing(True)
    L.execute(to_dev(raw), prof, hits)
    names = [n for n, _ in L.stage_times()]
    L.close()
    assert "stageC_colifft_detect_fold" in names, names
    chans = [0, 1, 45, 90, 126, 127]
    prof_o, hits_o = oracle_channels(oracle, raw, 8, chans, nchan, npol, ndat, cfreq, bw, dm, g["nfft"],
                                     g["impulse_pos"], g["impulse_neg"], dspx.STOKES, 4, nbin, pps, chunks, nsub)
    got = from_dev(prof)[:, chans]
    e = rel_rms(got, prof_o)
    print("profile %.3g" % e, flush=True)
    assert np.array_equal(from_dev(hits).view(np.uint32), hits_o)
    assert e < 1e-5


# beams of the uneven batch's draw (seed 3): 512 channels at nfft 128 and
# 256 (8- and 16-bit), 256 channels at 4096, 512 16-bit channels at 16384,
# 256 16-bit channels at 32768
SLICE = [36, 0, 8, 10, 2]


def _slice_beams(dspx):
    import importlib
    bench = importlib.import_module("bench")
    beams = bench.draw_uneven(dspx, 48, seed=3)
    return [beams[i] for i in SLICE], beams[24]


def _batch_case(ctx, dspx, oracle, beams, autotune, windows, cap_bytes=128 << 20):
    """Each beam cut to `windows` windows of four times its reference
    length (room for every length the autotuner tries) or to cap_bytes of
    raw input, but at least one reference window, through one batch; the
    oracle on the first, a middle and the last channel of each."""
    import torch
    descs, raws, folds, cut = [], [], [], []
    for i, b in enumerate(beams):
        g = b["geom"]
        per_sample = b["nchan"] * 2 * 2 * b["nbit"] // 8
        ndat = min(windows * 4 * g["nfft"] + g["noverlap"], cap_bytes // per_sample)
        ndat = min(b["ndat"], max(ndat, g["nfft"] + g["noverlap"]))
        cut.append(ndat)
        descs.append(dspx.ltf_desc(b["nbit"], b["nchan"], 2, ndat, b["cfreq"], b["bw"], b["dm"], b["nbin"],
                                   state=dspx.STOKES, ndim=4))
        raws.append(_noise(ndat, b["nchan"], 2, b["nbit"], seed=100 + i))
    B = dspx.Batch(ctx, descs, autotune=autotune)
    profs, hits = [], []
    for i, b in enumerate(beams):
        info = B.beams[i]
        rate = b["bw"] / b["nchan"] * 1e6
        chunks = dspx.fold_chunks(info["ndat_out"], rate, b["period"], info["nfilt_pos"], nsub=2)
        pps = (1.0 / rate) / b["period"]
        B.set_fold(i, pps, chunks)
        folds.append((pps, chunks))
        profs.append(torch.zeros((2, b["nchan"], info["nval"] * b["nbin"]), dtype=torch.float32, device="cuda"))
        hits.append(torch.zeros((2, b["nbin"]), dtype=torch.int32, device="cuda"))
    print("config uneven slice", [(B.beams[i]["nfft"], B.beams[i]["form"]) for i in range(len(beams))], flush=True)
    B.execute([to_dev(r) for r in raws], profs, hits)
    for i, b in enumerate(beams):
        info = B.beams[i]
        chans = [0, b["nchan"] // 2, b["nchan"] - 1]
        pps, chunks = folds[i]
        prof_o, hits_o = oracle_channels(oracle, raws[i], b["nbit"], chans, b["nchan"], 2, cut[i], b["cfreq"],
                                         b["bw"], b["dm"], info["nfft"], info["nfilt_pos"], info["nfilt_neg"],
                                         dspx.STOKES, 4, b["nbin"], pps, chunks, 2)
        got = from_dev(profs[i]).reshape(2, b["nchan"], 1, -1)[:, chans]
        e = rel_rms(got, prof_o)
        print("beam %d nfft %d: profile %.3g" % (SLICE[i] if i < len(SLICE) else -1, info["nfft"], e), flush=True)
        assert np.array_equal(from_dev(hits[i]).view(np.uint32), hits_o)
        assert e < 1e-5
    B.close()
    return B


def test_config_uneven_slice(ctx, dspx, oracle):
    """Five beams of the draw, autotuned as the benchmark runs them."""
    beams, _ = _slice_beams(dspx)
    _batch_case(ctx, dspx, oracle, beams, True, 3)


def test_config_uneven_longest_beam(ctx, dspx, oracle):
    """The draw's longest transform: 64 channels at nfft 2^21 (overlap
    153332), one window."""
    _, big = _slice_beams(dspx)
    assert big["geom"]["nfft"] == 1 << 21
    b = dict(big)
    g = b["geom"]
    b["ndat"] = g["nfft"] + 1000
    import torch
    desc = dspx.ltf_desc(b["nbit"], b["nchan"], 2, b["ndat"], b["cfreq"], b["bw"], b["dm"], b["nbin"],
                         state=dspx.STOKES, ndim=4)
    B = dspx.Batch(ctx, [desc], autotune=False)
    info = B.beams[0]
    assert info["nfft"] == 1 << 21
    raw = _noise(b["ndat"], b["nchan"], 2, b["nbit"], seed=24)
    rate = b["bw"] / b["nchan"] * 1e6
    chunks = dspx.fold_chunks(info["ndat_out"], rate, b["period"], info["nfilt_pos"], nsub=2)
    pps = (1.0 / rate) / b["period"]
    B.set_fold(0, pps, chunks)
    prof = torch.zeros((2, b["nchan"], info["nval"] * b["nbin"]), dtype=torch.float32, device="cuda")
    hits = torch.zeros((2, b["nbin"]), dtype=torch.int32, device="cuda")
    print("config longest beam", info, flush=True)
    B.execute([to_dev(raw)], [prof], [hits])
    chans = [0, 63]
    prof_o, hits_o = oracle_channels(oracle, raw, b["nbit"], chans, b["nchan"], 2, b["ndat"], b["cfreq"], b["bw"],
                                     b["dm"], info["nfft"], info["nfilt_pos"], info["nfilt_neg"], dspx.STOKES, 4,
                                     b["nbin"], pps, chunks, 2)
    got = from_dev(prof).reshape(2, b["nchan"], 1, -1)[:, chans]
    e = rel_rms(got, prof_o)
    print("profile %.3g" % e, flush=True)
    B.close()
    assert np.array_equal(from_dev(hits).view(np.uint32), hits_o)
    assert e < 1e-5
