"""Shared test helpers (TukeyResponse, error measures, device transfer)."""
import math

import numpy as np


def rel_rms(a, b):
    a = np.asarray(a).astype(np.complex128 if np.iscomplexobj(a) else np.float64).ravel()
    b = np.asarray(b).astype(np.complex128 if np.iscomplexobj(b) else np.float64).ravel()
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / np.sum(np.abs(b) ** 2)))


def tukey(size, transition_start, transition_end):
    """Apodization::Tukey (Signal/General/Apodization.C:123-146), float."""
    buf = np.empty(size, np.float32)
    ds = np.float32(2 * transition_start)
    de = np.float32(2 * transition_end)
    for i in range(size):
        if i < transition_start:
            buf[i] = 0.5 * (1 - math.cos(2.0 * math.pi * float(np.float32(i) / ds)))
        elif i + 1 > size - transition_end:
            buf[i] = 0.5 * (1 - math.cos(2.0 * math.pi * float(np.float32(size - i - 1) / de)))
        else:
            buf[i] = 1.0
    return buf


def make_test_response(oracle, nfft, impulse_neg, impulse_pos):
    """TestResponse::build / create_response (Signal/General/testutils/
    TestResponse.cpp:30-165): a zero-mean, Tukey-windowed one-cycle chirp;
    returns (temporal [neg+pos] complex64, spectral [nfft] complex64 scaled by
    1/sqrt(nfft))."""
    size = impulse_neg + impulse_pos
    amp = tukey(size, impulse_neg, impulse_pos)
    phase = 2.0 * np.pi * np.arange(size) / size
    temporal = (amp * (np.cos(phase) + 1j * np.sin(phase)).astype(np.complex64)).astype(np.complex64)
    sum_temporal = np.sum(temporal.astype(np.complex128))
    correction = np.sin(0.5 * phase).astype(np.float32)
    sum_correction = float(np.sum(correction.astype(np.float64)))
    mean_offset = sum_temporal / sum_correction
    temporal = (temporal - (mean_offset * correction.astype(np.float64))).astype(np.complex64)
    padded = np.zeros(nfft, np.complex64)
    padded[:impulse_pos] = temporal[impulse_neg:impulse_neg + impulse_pos]
    padded[nfft - impulse_neg:] = temporal[:impulse_neg]
    spectral = oracle.fft(padded, -1)
    spectral = (spectral * np.float32(1.0 / math.sqrt(nfft))).astype(np.complex64)
    return temporal, spectral


def tukey_kat_input(nfilt, n_neg, n_pos, n_loop):
    """Impulse train of ConvolutionTest.test_operate_tukey
    (Signal/General/tests/ConvolutionTest.cpp:176-212)."""
    n_overlap = n_neg + n_pos
    n_keep = nfilt - n_overlap
    expected_ndat = n_keep * 2 * n_loop
    input_ndat = expected_ndat + n_overlap
    x = np.zeros((1, 2, input_ndat), np.complex64)
    base = [n_overlap, nfilt - n_pos]
    offsets = np.zeros((n_loop, 2), np.int64)
    for il in range(n_loop):
        for ip in range(2):
            off = base[ip] + il * 2 * (nfilt - n_overlap) + il + ip
            offsets[il, ip] = off
            x[0, ip, off] = 1.0
    return x, offsets, expected_ndat


def tukey_kat_check(result, temporal, offsets, n_overlap, threshold=5e-8):
    """The assertions of test_operate_tukey (ConvolutionTest.cpp:234-307).
    result: [2][ndat] complex64.  Returns (max mean_error, max off-pulse rms)."""
    n_loop = offsets.shape[0]
    worst_err, worst_rms = 0.0, 0.0
    for ip in range(2):
        r = result[ip]
        base_pow, base_cnt, bidx = 0.0, 0, 0
        for il in range(n_loop):
            off = offsets[il, ip] - n_overlap
            seg = r[off:off + n_overlap].astype(np.complex128)
            e = temporal.astype(np.complex128)
            power_diff = np.sum(np.abs(seg - e) ** 2)
            power = np.sum(np.abs(e) ** 2)
            mean_error = math.sqrt(power_diff / power / n_overlap)
            worst_err = max(worst_err, mean_error)
            end_idx = r.size if il + 1 == n_loop else offsets[il + 1, ip] - n_overlap
            if bidx < off:
                base_pow += float(np.sum(np.abs(r[bidx:off].astype(np.complex128)) ** 2))
                base_cnt += off - bidx
                bidx = off
            bidx += n_overlap
            if bidx < end_idx:
                base_pow += float(np.sum(np.abs(r[bidx:end_idx].astype(np.complex128)) ** 2))
                base_cnt += end_idx - bidx
                bidx = end_idx
        worst_rms = max(worst_rms, math.sqrt(base_pow / base_cnt))
    return worst_err, worst_rms


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def from_dev(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def oracle_load_to_fold(oracle, raw, nbit, nchan, npol, ndat, cfreq, bw, dm, geom, state, ndim, nbin,
                        pps, chunks, nsub, scale=1.0, mode=0):
    """The reference CPU chain on the oracle: unpack -> Convolution ->
    Detection -> Fold per chunk.  Returns (filterbank, prof, hits)."""
    if nbit == 8:
        x = oracle.unpack8(raw, nchan, npol, ndat, scale=scale)
    else:
        x = oracle.unpack16(raw, nchan, npol, ndat)
    x = np.ascontiguousarray(x).view(np.complex64).reshape(nchan, npol, ndat)
    resp = oracle.response(cfreq, bw, nchan, dm, geom["nfft"], mode)
    fb = oracle.convolve(x, resp, geom["impulse_pos"], geom["impulse_neg"])
    det = oracle.detect(fb, state, ndim)
    nout = fb.shape[2]
    npol_o = det.shape[1]
    nd = det.shape[2] // nout
    prof = np.zeros((nsub, nchan, npol_o, nbin * nd), np.float32)
    hits = np.zeros((nsub, 1, nbin), np.uint32)
    for (i0, n, phi, isub) in chunks:
        oracle.fold(det, nd, phi, pps, nbin, idat_start=i0, ndat_fold=n, prof=prof[isub], hits=hits[isub])
    return fb, prof, hits[:, 0]


def sk_rfi_input(nchan, npol, nblk, M, seed, tone_chan=None, impulse_blk=None, zero_blk=None):
    """Complex Gaussian noise [nchan][npol][nblk*M] (sigma 10 per component)
    with optional RFI: a constant-amplitude tone filling channel tone_chan (SK
    near 0), a broadband burst filling block impulse_blk in every channel
    (large, non-Gaussian power), and zeroed samples in block zero_blk of
    channel 0 (every other sample), for the zeroed-data counts."""
    rng = np.random.default_rng(seed)
    n = nblk * M
    x = (rng.standard_normal((nchan, npol, n)) + 1j * rng.standard_normal((nchan, npol, n))) * 10
    if tone_chan is not None:
        t = np.arange(n)
        x[tone_chan] += 80 * np.exp(2j * np.pi * 0.123 * t)
    if impulse_blk is not None:
        s = slice(impulse_blk * M, impulse_blk * M + M // 8)
        x[:, :, s] *= 30
    if zero_blk is not None:
        x[0, :, zero_blk * M:(zero_blk + 1) * M:2] = 0
    return x.astype(np.complex64)
