"""ctypes wrapper of the CPU oracle (oracle.c) and of the reference's own C
files compiled into oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke test of the build
entry point and the CPU-baseline leg of bench.py, as the reference results
are checked against.  The product library never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i16p = np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so (and _ref/libref.so when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    u, i, d, U = C.c_uint, C.c_int, C.c_double, C.c_uint64
    L.ox_delay_time.restype = d
    L.ox_delay_time.argtypes = [d, d]
    L.ox_smearing_samples.restype = u
    L.ox_smearing_samples.argtypes = [d, d, u, d, i, u]
    L.ox_optimal_fft_length.restype = C.c_int64
    L.ox_optimal_fft_length.argtypes = [U, U, u]
    L.ox_conv_geometry.restype = i
    L.ox_conv_geometry.argtypes = [d, d, u, d, u, u64p]
    L.ox_conv_npart.restype = U
    L.ox_conv_npart.argtypes = [U, U, U]
    L.ox_unpack8.restype = None
    L.ox_unpack8.argtypes = [u8p, u, u, U, i, C.c_float, f32p]
    L.ox_unpack16.restype = None
    L.ox_unpack16.argtypes = [i16p, u, u, U, i, f32p]
    L.ox_fft_f32.restype = None
    L.ox_fft_f32.argtypes = [U, f32p, i]
    L.ox_fft_f64.restype = None
    L.ox_fft_f64.argtypes = [U, f64p, i]
    L.ox_channel_centre.restype = d
    L.ox_channel_centre.argtypes = [d, d, u, i, u]
    L.ox_chirp_phase.restype = d
    L.ox_chirp_phase.argtypes = [d, d, d, d]
    L.ox_chirp_freq.restype = d
    L.ox_chirp_freq.argtypes = [d, u, u]
    L.ox_chirp_phase_delay.restype = d
    L.ox_chirp_phase_delay.argtypes = [d, d, d, d, d]
    L.ox_dedisp_response_f32.restype = None
    L.ox_dedisp_response_f32.argtypes = [d, d, u, i, d, u, i, i, f32p]
    L.ox_dedisp_response_f64.restype = None
    L.ox_dedisp_response_f64.argtypes = [d, d, u, i, d, u, i, f64p]
    L.ox_interchan_delay.restype = d
    L.ox_interchan_delay.argtypes = [d, d, u, d, d]
    L.ox_sample_delays.restype = i
    L.ox_sample_delays.argtypes = [d, d, u, i, d, i64p, f64p, i64p, C.POINTER(C.c_int64), C.POINTER(U)]
    for name, p in (("ox_convolve_f32", f32p), ("ox_convolve_f64", f64p)):
        fn = getattr(L, name)
        fn.restype = U
        fn.argtypes = [p, u, u, U, p, u, u, u, p, C.c_void_p]
    L.ox_detect_square_law.restype = None
    L.ox_detect_square_law.argtypes = [f32p, u, u, U, i, f32p]
    L.ox_detect_polarimetry.restype = None
    L.ox_detect_polarimetry.argtypes = [f32p, u, U, u, i, f32p]
    L.ox_fold.restype = U
    L.ox_fold.argtypes = [f32p, u, u, u, U, U, U, d, d, u, f32p, u32p, u, i,
                          C.c_void_p, U, u, u, C.c_void_p]
    L.ox_sk_compute.restype = U
    L.ox_sk_compute.argtypes = [f32p, u, u, U, u, u, i, f32p, u32p]
    L.ox_sk_blocks.restype = U
    L.ox_sk_blocks.argtypes = [U, u, u, C.POINTER(U)]
    L.ox_sk_detect_skfb.restype = U
    L.ox_sk_detect_skfb.argtypes = [f32p, u32p, U, u, u, u, f32p, f32p, u8p]
    L.ox_sk_detect.restype = U
    L.ox_sk_detect.argtypes = [f32p, u32p, U, u, u, u, u, i, u, f32p, f32p, f32p, d, C.c_void_p, C.c_void_p,
                               C.c_void_p, u, u, C.c_void_p, u8p]
    for name, p in (("ox_filterbank_f32", f32p), ("ox_filterbank_f64", f64p)):
        fn = getattr(L, name)
        fn.restype = U
        fn.argtypes = [p, u, u, U, u, u, u, u, p, p]
    L.ox_sk_mask.restype = U
    L.ox_sk_mask.argtypes = [f32p, u, u, U, u, u, u8p, U, f32p]
    for name, pt in (("ox_convolve_real_f32", f32p), ("ox_convolve_real_f64", f64p)):
        fn = getattr(L, name)
        fn.restype = U
        fn.argtypes = [pt, u, u, U, pt, u, u, u, pt]
    L.ox_sk_detect_res.restype = U
    L.ox_sk_detect_res.argtypes = [f32p, u32p, U, u, u, u, u32p, u32p, f32p, u8p, u, i, u, d, C.c_void_p,
                                   C.c_void_p, C.c_void_p, u, u8p]
    L.ox_sk_mask_res.restype = U
    L.ox_sk_mask_res.argtypes = [f32p, u, u, U, u, u, u, u, u8p, f32p]
    _lib = L
    return L


def ref():
    """The reference's own optimize_fft.c / stokes_detect.c / cross_detect.c,
    or None where oracle/_ref was not built."""
    global _ref
    if _ref is not None:
        return _ref
    path = os.path.join(_HERE, "_ref", "libref.so")
    if not os.path.exists(path):
        return None
    R = C.CDLL(path)
    R.optimal_fft_length.restype = C.c_int64
    R.optimal_fft_length.argtypes = [C.c_uint64, C.c_uint64, C.c_uint, C.c_char]
    for name in ("stokes_detect", "cross_detect"):
        fn = getattr(R, name)
        fn.restype = None
        fn.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                       C.c_void_p, C.c_void_p, C.c_uint]
    _ref = R
    return R


# ---------------------------------------------------------------- helpers


def geometry(cfreq, bw, nchan, dm, nfft=0):
    """dict(impulse_neg, impulse_pos, nfft, noverlap, step, nunsupported)."""
    out = np.zeros(6, np.uint64)
    rc = lib().ox_conv_geometry(cfreq, bw, nchan, dm, nfft, out)
    if rc != 0:
        raise ValueError("oracle geometry failed rc=%d" % rc)
    keys = ("impulse_neg", "impulse_pos", "nfft", "noverlap", "step", "nunsupported")
    return {k: int(v) for k, v in zip(keys, out)}


def unpack8(raw, nchan, npol, ndat, scale=1.0, offset_binary=False):
    out = np.empty(nchan * npol * ndat * 2, np.float32)
    lib().ox_unpack8(np.ascontiguousarray(raw).view(np.uint8).ravel(), nchan, npol,
                     ndat, int(offset_binary), scale, out)
    return out.reshape(nchan, npol, ndat, 2)


def unpack16(raw, nchan, npol, ndat, offset_binary=False):
    out = np.empty(nchan * npol * ndat * 2, np.float32)
    lib().ox_unpack16(np.ascontiguousarray(raw).view(np.int16).ravel(), nchan, npol,
                      ndat, int(offset_binary), out)
    return out.reshape(nchan, npol, ndat, 2)


def fft(x, sign=-1):
    """Unnormalised complex FFT of a complex64/complex128 vector."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.complex64:
        buf = x.view(np.float32).copy()
        lib().ox_fft_f32(x.size, buf, sign)
        return buf.view(np.complex64)
    buf = x.astype(np.complex128).view(np.float64).copy()
    lib().ox_fft_f64(x.size, buf, sign)
    return buf.view(np.complex128)


def response(cfreq, bw, nchan, dm, nfft, mode=0, dc_centred=False, interchan=False):
    """[nchan][nfft] complex response in FFT-bin order (normaliser included).
    mode 0: float phase (reference parity); 1: FP64 phase rounded to float;
    'f64': FP64 truth in double.  interchan: each channel's fractional
    inter-channel delay in the phase (dspsr -K)."""
    if mode == "f64":
        out = np.empty(nchan * nfft * 2, np.float64)
        lib().ox_dedisp_response_f64(cfreq, bw, nchan, int(dc_centred), dm, nfft, int(interchan), out)
        return out.view(np.complex128).reshape(nchan, nfft)
    out = np.empty(nchan * nfft * 2, np.float32)
    lib().ox_dedisp_response_f32(cfreq, bw, nchan, int(dc_centred), dm, nfft, mode, int(interchan), out)
    return out.view(np.complex64).reshape(nchan, nfft)


def response_real(cfreq, bw, nchan, dm, nfft, mode=0):
    """The response of real-sampled (Nyquist) input: frequency order (not
    swapped; real input is not dual-sideband, Observation.C:95-102), DC of
    the first channel zero (Response.C:418-420), normaliser 1/sqrt(nfft)."""
    r = response(cfreq, bw, nchan, dm, nfft, mode)
    r = np.roll(r, nfft // 2, axis=1)  # undo the dual-sideband swap
    r[0, 0] = 0
    return r


def convolve_real(x, resp, nfilt_pos, nfilt_neg):
    """Real-sampled overlap-save convolution.  x: [nchan][npol][ndat] float32
    or float64 real samples; resp [nchan][nfft] (response_real).  Returns the
    complex output [nchan][npol][npart*step]."""
    x = np.ascontiguousarray(x)
    nchan, npol, ndat = x.shape
    nfft = resp.shape[1]
    step = nfft - nfilt_pos - nfilt_neg
    npart = 0 if ndat < 2 * nfft else (ndat - 2 * (nfilt_pos + nfilt_neg)) // (2 * step)
    if x.dtype == np.float32:
        ft, fn, ct = np.float32, lib().ox_convolve_real_f32, np.complex64
    else:
        ft, fn, ct = np.float64, lib().ox_convolve_real_f64, np.complex128
        x = x.astype(np.float64)
    resp = np.ascontiguousarray(resp, ct)
    out = np.zeros(nchan * npol * npart * step * 2, ft)
    n = fn(x.ravel(), nchan, npol, ndat, resp.view(ft).ravel(), nfft, nfilt_pos, nfilt_neg, out)
    assert n == npart
    return out.view(ct).reshape(nchan, npol, npart * step)


def sample_delays(cfreq, bw, nchan, dm, dc_centred=False):
    """SampleDelay with the dispersion delay function (one span): per channel
    the integer and fractional delay in samples relative to the band centre,
    the applied input offset, the zero delay and the total delay."""
    delay = np.zeros(nchan, np.int64)
    frac = np.zeros(nchan, np.float64)
    applied = np.zeros(nchan, np.int64)
    zd, tot = C.c_int64(0), C.c_uint64(0)
    rc = lib().ox_sample_delays(cfreq, bw, nchan, int(dc_centred), dm, delay, frac, applied, C.byref(zd), C.byref(tot))
    if rc != 0:
        raise ValueError("negative applied delay")
    return {"delay": delay, "frac": frac, "applied": applied, "zero_delay": zd.value, "total": tot.value}


def retard(x, applied, ndat_out):
    """SampleDelay::transformation (SampleDelay.C:262-297): output sample i of
    channel c is input sample i + applied[c] (every pol).  x: [nchan][npol][ndat]."""
    out = np.empty(x.shape[:2] + (ndat_out,) + x.shape[3:], x.dtype)
    for c in range(x.shape[0]):
        out[c] = x[c, :, applied[c]:applied[c] + ndat_out]
    return out


def convolve(x, resp, nfilt_pos, nfilt_neg, zero_dm=False):
    """Overlap-save convolution.  x: [nchan][npol][ndat] complex64/128;
    resp: [nchan][nfft] of the same precision.  Returns out (and zdm)."""
    x = np.ascontiguousarray(x)
    nchan, npol, ndat = x.shape
    nfft = resp.shape[1]
    step = nfft - nfilt_pos - nfilt_neg
    npart = 0 if ndat < nfft else (ndat - nfilt_pos - nfilt_neg) // step
    if x.dtype == np.complex64:
        ft, fn = np.float32, lib().ox_convolve_f32
        resp = np.ascontiguousarray(resp, np.complex64)
    else:
        ft, fn = np.float64, lib().ox_convolve_f64
        x = x.astype(np.complex128)
        resp = np.ascontiguousarray(resp, np.complex128)
    out = np.zeros(nchan * npol * npart * step * 2, ft)
    zdm = np.zeros_like(out) if zero_dm else None
    fn(x.view(ft).ravel(), nchan, npol, ndat, resp.view(ft).ravel(), nfft,
       nfilt_pos, nfilt_neg, out, zdm.ctypes.data if zero_dm else None)
    cdt = np.complex64 if ft == np.float32 else np.complex128
    out = out.view(cdt).reshape(nchan, npol, npart * step)
    if zero_dm:
        return out, zdm.view(cdt).reshape(nchan, npol, npart * step)
    return out


INTENSITY, PPQQ, COHERENCE, STOKES = 0, 1, 2, 3


def detect(x, state, ndim=4):
    """x: [nchan][npol][ndat] complex64.  Returns the detected FPT array
    [nchan][npol_out][ndat*ndim_out] as the reference lays it out."""
    x = np.ascontiguousarray(x, np.complex64)
    nchan, npol, ndat = x.shape
    flat = x.view(np.float32).ravel()
    if state in (INTENSITY, PPQQ):
        onpol = npol if state == PPQQ else 1
        out = np.zeros(nchan * onpol * ndat, np.float32)
        lib().ox_detect_square_law(flat, nchan, npol, ndat, int(state == PPQQ), out)
        return out.reshape(nchan, onpol, ndat)
    out = np.zeros(nchan * 4 * ndat, np.float32)
    lib().ox_detect_polarimetry(flat, nchan, ndat, ndim, int(state == STOKES), out)
    return out.reshape(nchan, 4 // ndim, ndat * ndim)


def fold(data, ndim, phi, pps, nbin, idat_start=0, ndat_fold=None, prof=None,
         hits=None, zeroed=False, weights=None, ndatperweight=0, weight_idat=0,
         hits_nchan=1, want_binplan=False):
    """data: [nchan][npol][ndat*ndim] float32 (detected, FPT).  Accumulates
    into prof [nchan][npol][nbin*ndim] and hits [hits_nchan][nbin]."""
    data = np.ascontiguousarray(data, np.float32)
    nchan, npol, nf = data.shape
    ndat = nf // ndim
    if ndat_fold is None:
        ndat_fold = ndat - idat_start
    if prof is None:
        prof = np.zeros((nchan, npol, nbin * ndim), np.float32)
    if hits is None:
        hits = np.zeros((hits_nchan, nbin), np.uint32)
    binplan = np.zeros(max(ndat_fold, 1), np.uint32) if want_binplan else None
    w = None
    nweights = 0
    if weights is not None:
        w = np.ascontiguousarray(weights, np.uint16)
        nweights = w.size
    nfolded = lib().ox_fold(data.ravel(), nchan, npol, ndim, ndat, idat_start, ndat_fold,
                            phi, pps, nbin, prof.reshape(-1), hits.reshape(-1), hits_nchan,
                            int(zeroed), w.ctypes.data if w is not None else None, nweights,
                            ndatperweight, weight_idat,
                            binplan.ctypes.data if want_binplan else None)
    if want_binplan:
        return prof, hits, nfolded, binplan[:ndat_fold]
    return prof, hits, nfolded


def sk_blocks(ndat, M, noverlap=1):
    """(npart, ndat_out) of Resolution::prepare."""
    nout = C.c_uint64()
    npart = lib().ox_sk_blocks(ndat, M, noverlap, C.byref(nout))
    return npart, nout.value


def sk_compute(x, M, zeroed=False, noverlap=1):
    """x: [nchan][npol][ndat] complex64 -> sums [npart][nchan][npol][2] float32,
    counts [npart][nchan] uint32 (blocks of M every M / noverlap samples)."""
    x = np.ascontiguousarray(x, np.complex64)
    nchan, npol, ndat = x.shape
    npart, _ = sk_blocks(ndat, M, noverlap)
    sums = np.zeros(max(npart, 1) * nchan * npol * 2, np.float32)
    counts = np.zeros(max(npart, 1) * nchan, np.uint32)
    lib().ox_sk_compute(x.view(np.float32).ravel(), nchan, npol, ndat, M, noverlap, int(zeroed),
                        sums, counts)
    sums, counts = sums[:npart * nchan * npol * 2], counts[:npart * nchan]
    return sums.reshape(npart, nchan, npol, 2), counts.reshape(npart, nchan)


def sk_detect_skfb(sums, counts, lower, upper, Nd=1):
    npart, nchan, npol, _ = sums.shape
    mask = np.zeros(npart * nchan, np.uint8)
    n = lib().ox_sk_detect_skfb(np.ascontiguousarray(sums, np.float32).ravel(),
                                np.ascontiguousarray(counts, np.uint32).ravel(), npart,
                                nchan, npol, Nd, np.ascontiguousarray(lower, np.float32),
                                np.ascontiguousarray(upper, np.float32), mask)
    return mask.reshape(npart, nchan), int(n)


SKFB, FSCR, TSCR = 1, 2, 4


def sk_detect(sums, counts, M, lower, upper, symmetric, pcsd=3.0, flags=SKFB, Nd=1, omit_outliers=False,
              tscr_limits=None, noverlap=1, channels=None):
    """SpectralKurtosis::detect with one resolution.  tscr_limits: {total:
    (lower, upper)} for the channel totals above M that detect_tscr tests;
    channels: per-channel include flags of the averaging detectors (None:
    every channel).  Returns the mask [npart][nchan] and the (possibly
    modified, with omit_outliers) sums and counts."""
    sums = np.array(sums, np.float32, copy=True)
    counts = np.array(counts, np.uint32, copy=True)
    npart, nchan, npol, _ = sums.shape
    mask = np.zeros(npart * nchan, np.uint8)
    lim = tscr_limits or {}
    ch = None if channels is None else np.ascontiguousarray(channels, np.uint8)
    tt = np.array(sorted(lim), np.uint64)
    tl = np.array([lim[int(t)][0] for t in tt], np.float32)
    tu = np.array([lim[int(t)][1] for t in tt], np.float32)
    n = lib().ox_sk_detect(sums.reshape(-1), counts.reshape(-1), npart, nchan, npol, M, Nd, int(omit_outliers),
                           flags, np.ascontiguousarray(lower, np.float32), np.ascontiguousarray(upper, np.float32),
                           np.ascontiguousarray(symmetric, np.float32), pcsd, tt.ctypes.data, tl.ctypes.data,
                           tu.ctypes.data, tt.size, noverlap, ch.ctypes.data if ch is not None else None, mask)
    if n == 2 ** 64 - 1:
        raise ValueError("detect_tscr: a channel total in (M, 32768) has no thresholds")
    return mask.reshape(npart, nchan), sums, counts


def sk_mask(x, M, mask, noverlap=1):
    """x: [nchan][npol][ndat] complex64, mask [npart][nchan] of the sums'
    blocks -> the masked output, rows of ndat samples of which the first
    ndat_out (= npart*M with noverlap 1) are written and the rest left zero."""
    x = np.ascontiguousarray(x, np.complex64)
    nchan, npol, ndat = x.shape
    npart = mask.shape[0]
    out = np.zeros_like(x)
    lib().ox_sk_mask(x.view(np.float32).reshape(-1), nchan, npol, ndat, M, noverlap,
                     np.ascontiguousarray(mask, np.uint8).reshape(-1), npart, out.view(np.float32).reshape(-1))
    return out


def sk_detect_res(sums, counts, ndat, resolutions, flags=SKFB, Nd=1, omit_outliers=False, pcsd=3.0,
                  tscr_limits=None):
    """SpectralKurtosis::detect with several resolutions.  sums/counts: the
    finest resolution's [npart_0][nchan][npol][2] and [npart_0][nchan];
    resolutions: [(M, noverlap, lower, upper, symmetric, channels or None)]
    sorted by M (tables of M+1 by count, channels the include flags).
    Returns the mask [npart_0][nchan] and the sums and counts after the
    in-place tscrunches (the coarsest resolution's in the first rows)."""
    sums = np.array(sums, np.float32, copy=True)
    counts = np.array(counts, np.uint32, copy=True)
    npart, nchan, npol, _ = sums.shape
    Ms = np.array([r[0] for r in resolutions], np.uint32)
    novs = np.array([r[1] for r in resolutions], np.uint32)
    tables = np.concatenate([np.concatenate([np.asarray(r[2], np.float32), np.asarray(r[3], np.float32),
                                             np.asarray(r[4], np.float32)]) for r in resolutions])
    chans = np.concatenate([np.ones(nchan, np.uint8) if r[5] is None else np.asarray(r[5], np.uint8)
                            for r in resolutions])
    lim = tscr_limits or {}
    tt = np.array(sorted(lim), np.uint64)
    tl = np.array([lim[int(t)][0] for t in tt], np.float32)
    tu = np.array([lim[int(t)][1] for t in tt], np.float32)
    mask = np.zeros(npart * nchan, np.uint8)
    n = lib().ox_sk_detect_res(sums.reshape(-1), counts.reshape(-1), ndat, nchan, npol, len(resolutions), Ms, novs,
                               tables, chans, Nd, int(omit_outliers), flags, pcsd, tt.ctypes.data, tl.ctypes.data,
                               tu.ctypes.data, tt.size, mask)
    if n == 2 ** 64 - 1:
        raise ValueError("detect_tscr: a channel total in (M, 32768) has no thresholds")
    if n == 2 ** 64 - 2:
        raise ValueError("resolutions not compatible")
    return mask.reshape(npart, nchan), sums, counts


def sk_mask_res(x, M0, nov0, Ml, novl, mask):
    """SpectralKurtosis::mask with several resolutions (finest M0, noverlap
    nov0; coarsest Ml, novl) -> the masked output, rows of ndat samples of
    which the first ndat_out are written."""
    x = np.ascontiguousarray(x, np.complex64)
    nchan, npol, ndat = x.shape
    out = np.zeros_like(x)
    n = lib().ox_sk_mask_res(x.view(np.float32).reshape(-1), nchan, npol, ndat, M0, nov0, Ml, novl,
                             np.ascontiguousarray(mask, np.uint8).reshape(-1), out.view(np.float32).reshape(-1))
    return out, int(n)


def filterbank(x, nsub, fres, nfilt_pos, nfilt_neg, resp, f64=False):
    """FilterbankEngineCPU::perform: x [nchan][npol][ndat] complex, resp
    [nchan][nsub*fres] complex in application order -> out
    [nchan*nsub][npol][npart*nkeep] complex."""
    dt, ct = (np.float64, np.complex128) if f64 else (np.float32, np.complex64)
    x = np.ascontiguousarray(x, ct)
    nchan, npol, ndat = x.shape
    nfft = nsub * fres
    nkeep = fres - nfilt_pos - nfilt_neg
    novl = (nfilt_pos + nfilt_neg) * nsub
    npart = (ndat - novl) // (nfft - novl) if ndat > novl else 0
    out = np.zeros((nchan * nsub, npol, npart * nkeep), ct)
    fn = lib().ox_filterbank_f64 if f64 else lib().ox_filterbank_f32
    n = fn(x.view(dt).reshape(-1), nchan, npol, ndat, nsub, fres, nfilt_pos, nfilt_neg,
           np.ascontiguousarray(resp, ct).view(dt).reshape(-1), out.view(dt).reshape(-1))
    assert n == npart
    return out


def filterbank_response(cfreq, bw, nchan_in, nsub, dm, fres, mode=0):
    """The dedispersion response of a filterbank of nchan_in*nsub output
    channels of fres bins, in the order Response::swap_as_needed leaves it
    for single-channel complex input (Response.C:335-345: built per channel
    in frequency order, then the whole buffer's halves swapped), with the
    normaliser 1/sqrt(nsub*fres*fres) (Filterbank.C:176-185).  nchan_in must
    be 1 (the whole swap)."""
    assert nchan_in == 1
    nout = nchan_in * nsub
    r = response(cfreq, bw, nout, dm, fres, mode).reshape(nout, fres)  # per-channel swapped, 1/sqrt(fres)
    nat = np.roll(r, fres // 2, axis=1)  # undo the per-channel swap: frequency order within each channel
    flat = nat.reshape(-1) * np.float32(1.0 / np.sqrt(nsub * fres))
    return np.roll(flat, (nsub * fres) // 2).reshape(nchan_in, nsub * fres).astype(np.complex64)
