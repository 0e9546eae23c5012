"""dspx -- MI355X-native coherent dedispersion and folding for dspsr pipelines.

Python bindings of the C ABI in include/dspx.h (libdspx.so, built in this
directory by the Makefile).  Device buffers are plain device pointers; torch
tensors on a HIP device are accepted wherever a pointer is expected.

The compute path is the HIP library only: if libdspx.so is missing every call
raises, there is no CPU fallback.
"""
import atexit
import ctypes as C
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdspx.so")
_lib = None

OK, E_PARAM, E_STATE, E_HIP, E_UNSUPPORTED, E_INTERNAL = range(6)
CHIRP_TABLE, CHIRP_PARITY, CHIRP_TRUTH = 0, 1, 2
FORM_FUSED, FORM_PLAIN, FORM_TWOLEVEL, FORM_ONEPASS = 0, 1, 2, 3
INTENSITY, PPQQ, COHERENCE, STOKES = 0, 1, 2, 3


class DspxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dspx error %d: %s" % (code, msg))
        self.code = code


class DedispDesc(C.Structure):
    _fields_ = [("centre_frequency", C.c_double), ("bandwidth", C.c_double),
                ("dispersion_measure", C.c_double), ("nchan", C.c_uint32),
                ("dc_centred", C.c_uint32), ("nfft", C.c_uint32), ("interchan", C.c_uint32)]


class DadaInfo(C.Structure):
    _fields_ = [("centre_frequency", C.c_double), ("bandwidth", C.c_double), ("tsamp", C.c_double),
                ("hdr_size", C.c_uint64), ("obs_offset", C.c_uint64), ("ndat", C.c_uint64),
                ("picoseconds", C.c_int64), ("nchan", C.c_uint32), ("npol", C.c_uint32), ("nbit", C.c_uint32),
                ("ndim", C.c_uint32), ("resolution", C.c_uint32), ("reserved", C.c_uint32),
                ("utc_start", C.c_char * 32), ("source", C.c_char * 64), ("telescope", C.c_char * 64),
                ("instrument", C.c_char * 64)]


class Geometry(C.Structure):
    _fields_ = [("impulse_neg", C.c_uint32), ("impulse_pos", C.c_uint32), ("nfft", C.c_uint32),
                ("nsamp_overlap", C.c_uint32), ("nsamp_step", C.c_uint32),
                ("nunsupported", C.c_uint32)]


class ConvDesc(C.Structure):
    _fields_ = [("dedisp", DedispDesc), ("nchan", C.c_uint32), ("npol", C.c_uint32),
                ("nfft", C.c_uint32), ("nfilt_pos", C.c_uint32), ("nfilt_neg", C.c_uint32),
                ("chirp_mode", C.c_uint32), ("form", C.c_uint32), ("nyquist", C.c_uint32)]


class FoldChunk(C.Structure):
    _fields_ = [("idat_start", C.c_uint64), ("ndat", C.c_uint64), ("phi", C.c_double),
                ("isub", C.c_uint32), ("reserved", C.c_uint32)]


class LtfDesc(C.Structure):
    _fields_ = [("nbit", C.c_uint32), ("offset_binary", C.c_uint32), ("scale", C.c_float),
                ("nchan", C.c_uint32), ("npol", C.c_uint32), ("reserved0", C.c_uint32),
                ("ndat", C.c_uint64), ("dedisp", DedispDesc), ("nfft", C.c_uint32),
                ("nfilt_pos", C.c_uint32), ("nfilt_neg", C.c_uint32), ("chirp_mode", C.c_uint32),
                ("state", C.c_uint32), ("ndim", C.c_uint32), ("nbin", C.c_uint32),
                ("form", C.c_uint32)]


class FoldOpts(C.Structure):
    _fields_ = [("weights", C.c_void_p), ("nweights", C.c_uint64), ("ndatperweight", C.c_uint32),
                ("weight_idat", C.c_uint32), ("zeroed_samples", C.c_uint32), ("reserved", C.c_uint32)]


class FbDesc(C.Structure):
    _fields_ = [("nchan", C.c_uint32), ("npol", C.c_uint32), ("nsub", C.c_uint32), ("fres", C.c_uint32),
                ("nfilt_pos", C.c_uint32), ("nfilt_neg", C.c_uint32), ("reserved", C.c_uint32 * 2)]


class SkDesc(C.Structure):
    _fields_ = [("nchan", C.c_uint32), ("npol", C.c_uint32), ("M", C.c_uint32), ("Nd", C.c_uint32),
                ("omit_outliers", C.c_uint32), ("noverlap", C.c_uint32)]


SK_SKFB, SK_FSCR, SK_TSCR = 1, 2, 4


class Mjd(C.Structure):
    _fields_ = [("day", C.c_int64), ("sec", C.c_int64), ("frac", C.c_double)]


class TimedivBounds(C.Structure):
    _fields_ = [("idat_start", C.c_uint64), ("ndat", C.c_uint64), ("division", C.c_uint64),
                ("division_ndat", C.c_uint64), ("new_division", C.c_int32), ("end_reached", C.c_int32),
                ("in_next", C.c_int32), ("is_valid", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

P = C.c_void_p
U32, U64, I32, D = C.c_uint32, C.c_uint64, C.c_int, C.c_double

_SIGS = {
    "dspx_version": (C.c_char_p, []),
    "dspx_ctx_create": (I32, [I32, P, C.POINTER(P)]),
    "dspx_ctx_destroy": (None, [P]),
    "dspx_last_error": (C.c_char_p, [P]),
    "dspx_ctx_synchronize": (I32, [P]),
    "dspx_ctx_stream": (P, [P]),
    "dspx_dedisp_geometry": (I32, [C.POINTER(DedispDesc), C.POINTER(Geometry)]),
    "dspx_sample_delays": (I32, [C.POINTER(DedispDesc), P, P, P, C.POINTER(C.c_int64), C.POINTER(U64)]),
    "dspx_ascii_header_get": (I32, [C.c_char_p, C.c_char_p, C.c_char_p, U32]),
    "dspx_dada_parse": (I32, [C.c_char_p, C.POINTER(DadaInfo), C.c_char_p]),
    "dspx_delay_create": (I32, [P, U32, U32, P, C.POINTER(P)]),
    "dspx_delay_destroy": (None, [P]),
    "dspx_delay_retard": (I32, [P, P, U64, U64, U32, U64, P, U64, U64]),
    "dspx_unpack": (I32, [P, P, U32, U32, U32, U32, U64, C.c_float, P, U64, U64]),
    "dspx_conv_create": (I32, [P, C.POINTER(ConvDesc), C.POINTER(P)]),
    "dspx_conv_destroy": (None, [P]),
    "dspx_conv_set_response": (I32, [P, P]),
    "dspx_conv_perform": (I32, [P, P, U64, U64, U64, P, U64, U64, P, U32]),
    "dspx_fb_create": (I32, [P, C.POINTER(FbDesc), C.POINTER(P)]),
    "dspx_fb_destroy": (None, [P]),
    "dspx_fb_set_response": (I32, [P, P]),
    "dspx_fb_geometry": (I32, [P, U64, C.POINTER(U64), C.POINTER(U64)]),
    "dspx_fb_perform": (I32, [P, P, U64, U64, U64, P, U64, U64]),
    "dspx_fb_perform_zdm": (I32, [P, P, U64, U64, U64, P, U64, U64, P]),
    "dspx_fb_batches": (I32, [P, C.POINTER(U32)]),
    "dspx_detect": (I32, [P, U32, U32, P, U64, U64, U32, U32, U64, P, U64, U64]),
    "dspx_fold_create": (I32, [P, U64, D, U32, C.POINTER(FoldChunk), U32, C.POINTER(P)]),
    "dspx_fold_destroy": (None, [P]),
    "dspx_fold_binplan": (I32, [P, P]),
    "dspx_fold_binplan_host": (I32, [D, D, U32, U64, P]),
    "dspx_fold_nsegments_host": (I32, [D, D, U64, P]),
    "dspx_fold_hits": (I32, [P, P]),
    "dspx_fold_execute": (I32, [P, P, U64, U64, U32, U32, U32, P]),
    "dspx_fold_create_opts": (I32, [P, U64, D, U32, C.POINTER(FoldChunk), U32, C.POINTER(FoldOpts), C.POINTER(P)]),
    "dspx_fold_replan": (I32, [P, U64, D, U32, C.POINTER(FoldChunk), U32, C.POINTER(FoldOpts)]),
    "dspx_fold_ndat_folded": (I32, [P, C.POINTER(U64)]),
    "dspx_fold_execute_zeroed": (I32, [P, P, U64, U64, U32, U32, U32, P, P, U32, P]),
    "dspx_fold_static": (I32, [P, U64, D, U32, C.POINTER(FoldChunk), U32, P, U64, U64, U32, U32, U32, P, P]),
    "dspx_sk_create": (I32, [P, C.POINTER(SkDesc), C.POINTER(P)]),
    "dspx_sk_destroy": (None, [P]),
    "dspx_sk_set_limits": (I32, [P, P, P, P, D]),
    "dspx_sk_compute": (I32, [P, P, U64, U64, U64, I32, P, P]),
    "dspx_sk_compute_raw": (I32, [P, P, U32, U32, C.c_float, U64, U64, P, P]),
    "dspx_sk_detect": (I32, [P, P, P, U64, U32, P]),
    "dspx_sk_tscr_totals": (I32, [P, P, U64, P]),
    "dspx_sk_set_channels": (I32, [P, P]),
    "dspx_sk_blocks": (I32, [P, U64, C.POINTER(U64), C.POINTER(U64)]),
    "dspx_sk_set_tscr_limits": (I32, [P, U32, P, P, P]),
    "dspx_sk_detect_tscr": (I32, [P, P, P, U64, P]),
    "dspx_sk_mask": (I32, [P, P, P, U64, U64, P, U64, U64, U64]),
    "dspx_sk_add_resolution": (I32, [P, U32, U32, P, P, P, P]),
    "dspx_sk_clear_resolutions": (I32, [P]),
    "dspx_sk_blocks_res": (I32, [P, U64, P, P]),
    "dspx_sk_detect_res": (I32, [P, P, P, U64, U32, P]),
    "dspx_sk_res_totals": (I32, [P, P]),
    "dspx_sk_res_tscr": (I32, [P, P]),
    "dspx_sk_mask_res": (I32, [P, P, P, U64, U64, P, U64, U64, U64]),
    "dspx_batch_create": (I32, [P, C.POINTER(LtfDesc), U32, U32, C.POINTER(P)]),
    "dspx_batch_destroy": (None, [P]),
    "dspx_batch_beam": (I32, [P, U32, C.POINTER(LtfDesc), C.POINTER(U64), C.POINTER(U64)]),
    "dspx_batch_set_fold": (I32, [P, U32, D, C.POINTER(FoldChunk), U32]),
    "dspx_batch_execute": (I32, [P, P, P, P]),
    "dspx_batch_capture": (I32, [P, P, P, P]),
    "dspx_batch_replay": (I32, [P]),
    "dspx_rfinet_nparam": (U32, []),
    "dspx_rfinet_create": (I32, [P, U64, C.POINTER(P)]),
    "dspx_rfinet_destroy": (None, [P]),
    "dspx_rfinet_get_params": (I32, [P, P, I32]),
    "dspx_rfinet_set_params": (I32, [P, P]),
    "dspx_rfi_features": (I32, [P, P, P, U64, U32, U32, U32, P, P]),
    "dspx_rfinet_forward": (I32, [P, P, U64, U32, P]),
    "dspx_rfinet_backward": (I32, [P, P, P, U64, U32, C.c_float, P]),
    "dspx_rfinet_grad": (I32, [P, P]),
    "dspx_rfinet_adam": (I32, [P, C.c_float, C.c_float, C.c_float, C.c_float, U32]),
    "dspx_rfinet_propose": (I32, [P, P, U64, C.c_float, P]),
    "dspx_ltf_create": (I32, [P, C.POINTER(LtfDesc), C.POINTER(P)]),
    "dspx_ltf_destroy": (None, [P]),
    "dspx_ltf_geometry": (I32, [P, C.POINTER(U64), C.POINTER(U64)]),
    "dspx_ltf_set_fold": (I32, [P, D, C.POINTER(FoldChunk), U32]),
    "dspx_timediv_create": (I32, [P, D, D, D, P, C.POINTER(P)]),
    "dspx_timediv_destroy": (None, [P]),
    "dspx_timediv_set_bounds": (I32, [P, P, D, U64, P]),
    "dspx_timediv_block": (I32, [P, P, D, U64, P, U32, C.POINTER(U32)]),
    "dspx_ltf_execute": (I32, [P, P, P, P]),
    "dspx_ltf_execute_masked": (I32, [P, P, P, U32, U64, P, P]),
    "dspx_ltf_filterbank": (I32, [P, P, P]),
    "dspx_ltf_stream_reset": (I32, [P]),
    "dspx_ltf_stream_next": (I32, [P, C.POINTER(U64), C.POINTER(U64), C.POINTER(U64)]),
    "dspx_ltf_stream_execute": (I32, [P, P, P, P]),
    "dspx_ltf_stream_filterbank": (I32, [P, P, P]),
    "dspx_ltf_set_timing": (I32, [P, I32]),
    "dspx_ltf_stage_times": (I32, [P, P, P, I32, C.POINTER(I32)]),
}


# Live handles, closed at interpreter exit (engines before their contexts),
# while the HIP runtime and the streams a context was given still exist:
# finalizers that run during interpreter teardown may come after the runtime
# is gone.
_live_engines = weakref.WeakSet()
_live_contexts = weakref.WeakSet()
_atexit_registered = False


def _close_all():
    for o in list(_live_engines):
        o.close()
    for o in list(_live_contexts):
        o.close()


def _register(obj, is_context=False):
    global _atexit_registered
    if not _atexit_registered:
        # registered when the first handle is made (after torch has set up
        # the runtime), so it runs before torch's own exit handlers
        atexit.register(_close_all)
        _atexit_registered = True
    (_live_contexts if is_context else _live_engines).add(obj)


def lib():
    """The loaded libdspx.so; raises when the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdspx.so not built (run make -C %s)" % _HERE)
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return sorted(_SIGS)


def ptr(x):
    """Device pointer of a torch tensor, or an int / None passed through."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _check(rc, ctx=None):
    if rc != OK:
        msg = lib().dspx_last_error(ctx).decode() if ctx else "(no context)"
        raise DspxError(rc, msg)


def geometry(centre_frequency, bandwidth, nchan, dm, nfft=0, dc_centred=False):
    """Convolution geometry by the reference rule (host only)."""
    d = DedispDesc(centre_frequency, bandwidth, dm, nchan, int(dc_centred), nfft, 0)
    g = Geometry()
    rc = lib().dspx_dedisp_geometry(C.byref(d), C.byref(g))
    if rc != OK:
        raise DspxError(rc, "geometry: smearing exceeds the limit or nfft too small")
    return {"impulse_neg": g.impulse_neg, "impulse_pos": g.impulse_pos, "nfft": g.nfft,
            "noverlap": g.nsamp_overlap, "step": g.nsamp_step, "nunsupported": g.nunsupported}


class TimeDivide:
    """dspx_timediv: sub-integration bounds across blocks (TimeDivide,
    TimeDivide.C:36-506) for divisions of `seconds`, or of `turns` of a
    constant folding period.  Epochs are (day, second of the day, fraction
    of a second) triples."""

    def __init__(self, start, seconds=0.0, turns=0.0, period=0.0, reference_epoch=None):
        self.h = None
        h = P()
        ref = C.byref(Mjd(*reference_epoch)) if reference_epoch is not None else None
        _check(lib().dspx_timediv_create(C.byref(Mjd(*start)), seconds, turns, period, ref, C.byref(h)))
        self.h = h
        self._lib = lib()

    def close(self):
        if self.h:
            self._lib.dspx_timediv_destroy(self.h)
            self.h = None

    __del__ = close

    def set_bounds(self, start, rate, ndat):
        b = TimedivBounds()
        _check(lib().dspx_timediv_set_bounds(self.h, C.byref(Mjd(*start)), rate, ndat, C.byref(b)))
        return b.as_dict()

    def block(self, start, rate, ndat, max_div=4096):
        """The set_bounds calls Subint::transformation makes for one block:
        the valid bounds, in order."""
        arr = (TimedivBounds * max_div)()
        n = C.c_uint32()
        _check(lib().dspx_timediv_block(self.h, C.byref(Mjd(*start)), rate, ndat, arr, max_div, C.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def fold_chunks(self, start, rate, ndat, period, reference_phase=0.0, t_ref=(0, 0, 0.0)):
        """The block's bounds as fold chunks [(idat_start, ndat, phi,
        division)], phi the phase of each chunk's first sample (constant
        period, zero phase at t_ref plus reference_phase)."""
        out = []
        for b in self.block(start, rate, ndat):
            day, sec, frac = start
            t = (float(day - t_ref[0]) * 86400.0 + float(sec - t_ref[1]) + (frac - t_ref[2])
                 + (b["idat_start"] + 0.5) / rate)
            out.append((b["idat_start"], b["ndat"], fold_phi(t, period, reference_phase), b["division"]))
        return out


def binplan_host(phi, pps, nbin, ndat):
    import numpy as np
    out = np.zeros(max(ndat, 1), np.uint32)
    _check(lib().dspx_fold_binplan_host(phi, pps, nbin, ndat, out.ctypes.data))
    return out[:ndat]


class Context:
    """dspx_ctx.  ``stream``: a HIP stream handle to enqueue on; None or 0
    makes the library create a stream of its own (then torch work on the
    default stream is NOT ordered with the library's)."""

    def __init__(self, device=0, stream=None):
        self.h = None
        self._lib = lib()  # held so that close() still works at interpreter exit
        h = P()
        _check(lib().dspx_ctx_create(device, stream, C.byref(h)))
        self.h = h
        _register(self, True)

    def close(self):
        if self.h:
            # engines made on this context hold its stream and device: they
            # are destroyed first, never after the context is gone
            for o in list(_live_engines):
                if getattr(o, "ctx", None) is self:
                    o.close()
            self._lib.dspx_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def check(self, rc):
        _check(rc, self.h)

    def synchronize(self):
        self.check(lib().dspx_ctx_synchronize(self.h))

    @property
    def stream(self):
        return lib().dspx_ctx_stream(self.h)

    # ---- engines
    def unpack(self, raw, nbit, nchan, npol, ndat, out, chan_stride, pol_stride, scale=1.0,
               offset_binary=False):
        self.check(lib().dspx_unpack(self.h, ptr(raw), nbit, int(offset_binary), nchan, npol, ndat,
                                     scale, ptr(out), chan_stride, pol_stride))

    def detect(self, state, ndim, x, in_chan_stride, in_pol_stride, nchan, npol, ndat, out,
               out_chan_stride, out_pol_stride):
        self.check(lib().dspx_detect(self.h, state, ndim, ptr(x), in_chan_stride, in_pol_stride,
                                     nchan, npol, ndat, ptr(out), out_chan_stride, out_pol_stride))


def _dedisp(centre_frequency, bandwidth, dm, nchan, dc_centred=False, interchan=False):
    return DedispDesc(centre_frequency, bandwidth, dm, nchan, int(dc_centred), 0, int(interchan))


def sample_delays(centre_frequency, bandwidth, nchan, dm, dc_centred=False):
    """SampleDelay with the dispersion delay function (host only): per channel
    the integer and fractional delays relative to the band centre and the
    applied input offset; the zero delay and the total delay."""
    import numpy as np
    d = _dedisp(centre_frequency, bandwidth, dm, nchan, dc_centred, True)
    delay = np.zeros(nchan, np.int64)
    frac = np.zeros(nchan, np.float64)
    applied = np.zeros(nchan, np.int64)
    zd, tot = C.c_int64(0), U64(0)
    rc = lib().dspx_sample_delays(C.byref(d), delay.ctypes.data, frac.ctypes.data, applied.ctypes.data,
                                  C.byref(zd), C.byref(tot))
    if rc != OK:
        raise DspxError(rc, "sample_delays: invalid band or a negative offset")
    return {"delay": delay, "frac": frac, "applied": applied, "zero_delay": zd.value, "total": tot.value}


def ascii_header_get(header, key):
    """The value token of key in a DADA ASCII header, or None."""
    buf = C.create_string_buffer(4096)
    h = header.encode() if isinstance(header, str) else bytes(header)
    rc = lib().dspx_ascii_header_get(h, key.encode(), buf, len(buf))
    return buf.value.decode() if rc == 1 else None


def dada_parse(header):
    """DADA header fields (ASCIIObservation::load) as a dict; raises
    DspxError naming a missing required key."""
    h = header.encode() if isinstance(header, str) else bytes(header)
    info = DadaInfo()
    missing = C.create_string_buffer(32)
    rc = lib().dspx_dada_parse(h, C.byref(info), missing)
    if rc != OK:
        raise DspxError(rc, "DADA header: %s" % (("missing or invalid " + missing.value.decode())
                                                 if missing.value else "invalid NDIM"))
    out = {k: getattr(info, k) for k, _ in DadaInfo._fields_ if k != "reserved"}
    for k in ("utc_start", "source", "telescope", "instrument"):
        out[k] = out[k].decode()
    return out


def read_dada(path, default_header_size=4096):
    """DADAFile: the header (DADAHeader::load_header: HDR_SIZE bytes, re-read
    when HDR_SIZE exceeds the default) and the raw samples after it as a
    read-only memory map [ndat][nchan][npol][ndim * nbit / 8] of bytes
    (TFP), ndat limited by NDAT when the header gives it."""
    import numpy as np
    size = default_header_size
    while True:
        with open(path, "rb") as f:
            head = f.read(size)
        if len(head) < size:
            raise DspxError(E_PARAM, "%s: shorter than its %d-byte header" % (path, size))
        text = head.split(b"\0", 1)[0]
        hs = ascii_header_get(text, "HDR_SIZE")
        hdr = int(hs) if hs else 0
        if hdr <= size:
            break
        size = hdr
    if hdr == 0:
        raise DspxError(E_PARAM, "%s: HDR_SIZE undefined" % path)
    info = dada_parse(text)
    bps = info["nchan"] * info["npol"] * info["ndim"] * info["nbit"] // 8
    nbytes = os.path.getsize(path) - hdr
    ndat = nbytes // bps
    if info["ndat"]:
        ndat = min(ndat, info["ndat"])
    raw = np.memmap(path, dtype=np.uint8, mode="r", offset=hdr,
                    shape=(ndat, info["nchan"], info["npol"], info["ndim"] * info["nbit"] // 8))
    return info, raw


class SampleDelay:
    """SampleDelay::Engine: fixed per-(channel, pol) input offsets."""

    def __init__(self, ctx, nchan, npol, applied):
        import numpy as np
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        a = np.ascontiguousarray(np.asarray(applied, np.int64).reshape(nchan, -1))
        if a.shape[1] == 1 and npol > 1:
            a = np.ascontiguousarray(np.repeat(a, npol, axis=1))
        h = P()
        ctx.check(lib().dspx_delay_create(ctx.h, nchan, npol, a.ctypes.data, C.byref(h)))
        self.h = h
        _register(self, False)

    def close(self):
        if self.h:
            self._lib.dspx_delay_destroy(self.h)
            self.h = None

    __del__ = close

    def retard(self, x, in_chan_stride, in_pol_stride, ndim, ndat_out, out, out_chan_stride, out_pol_stride):
        self.ctx.check(lib().dspx_delay_retard(self.h, ptr(x), in_chan_stride, in_pol_stride, ndim, ndat_out,
                                               ptr(out), out_chan_stride, out_pol_stride))


class Convolution:
    """Convolution::Engine (fused or plain form)."""

    def __init__(self, ctx, nchan, npol, nfft, nfilt_pos, nfilt_neg, chirp_mode=CHIRP_PARITY,
                 form=FORM_FUSED, centre_frequency=0.0, bandwidth=0.0, dm=0.0, dc_centred=False,
                 interchan=False, nyquist=False):
        """nyquist: real-sampled input (ndat real samples per chan/pol; the
        complex output at half the rate)."""
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        d = ConvDesc(_dedisp(centre_frequency, bandwidth, dm, nchan, dc_centred, interchan), nchan, npol, nfft,
                     nfilt_pos, nfilt_neg, chirp_mode, form, int(nyquist))
        h = P()
        ctx.check(lib().dspx_conv_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        _register(self, False)
        self.step = nfft - nfilt_pos - nfilt_neg

    def close(self):
        if self.h:
            self._lib.dspx_conv_destroy(self.h)
            self.h = None

    __del__ = close

    def set_response(self, resp):
        import numpy as np
        r = np.ascontiguousarray(resp, np.complex64)
        self.ctx.check(lib().dspx_conv_set_response(self.h, r.ctypes.data))

    def perform(self, x, in_chan_stride, in_pol_stride, ndat, out, out_chan_stride, out_pol_stride,
                npart, zero_dm=None):
        self.ctx.check(lib().dspx_conv_perform(self.h, ptr(x), in_chan_stride, in_pol_stride, ndat,
                                               ptr(out), out_chan_stride, out_pol_stride,
                                               ptr(zero_dm), npart))


class Fold:
    """Fold::Engine: device bin plan + accumulation.  weights (uint16 array,
    with ndatperweight and weight_idat) and zeroed follow Fold::fold's
    WeightedTimeSeries and zeroed-data options (Fold.C:700-733, 893-930)."""

    def __init__(self, ctx, ndat, pps, nbin, chunks, weights=None, ndatperweight=0, weight_idat=0,
                 zeroed=False):
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        arr, opts, keep = self._args(chunks, weights, ndatperweight, weight_idat, zeroed)
        h = P()
        ctx.check(lib().dspx_fold_create_opts(ctx.h, ndat, pps, nbin, arr, len(chunks), C.byref(opts), C.byref(h)))
        self.h = h
        _register(self, False)

    @staticmethod
    def _args(chunks, weights, ndatperweight, weight_idat, zeroed):
        import numpy as np
        arr = (FoldChunk * len(chunks))(*[FoldChunk(c[0], c[1], c[2], c[3], 0) for c in chunks])
        w = None if weights is None else np.ascontiguousarray(weights, np.uint16)
        opts = FoldOpts(w.ctypes.data if w is not None else None, 0 if w is None else w.size,
                        ndatperweight if w is not None else 0, weight_idat, int(bool(zeroed)), 0)
        return arr, opts, w

    def replan(self, ndat, pps, nbin, chunks, weights=None, ndatperweight=0, weight_idat=0, zeroed=False):
        arr, opts, keep = self._args(chunks, weights, ndatperweight, weight_idat, zeroed)
        self.ctx.check(lib().dspx_fold_replan(self.h, ndat, pps, nbin, arr, len(chunks), C.byref(opts)))

    def ndat_folded(self):
        n = U64()
        _check(lib().dspx_fold_ndat_folded(self.h, C.byref(n)))
        return n.value

    def close(self):
        if self.h:
            self._lib.dspx_fold_destroy(self.h)
            self.h = None

    __del__ = close

    def binplan(self, out):
        self.ctx.check(lib().dspx_fold_binplan(self.h, ptr(out)))

    def hits(self, out):
        self.ctx.check(lib().dspx_fold_hits(self.h, ptr(out)))

    def execute(self, x, chan_stride, pol_stride, nchan, npol, ndim, prof):
        self.ctx.check(lib().dspx_fold_execute(self.h, ptr(x), chan_stride, pol_stride, nchan, npol,
                                               ndim, ptr(prof)))

    def execute_zeroed(self, x, chan_stride, pol_stride, nchan, npol, ndim, prof, hits, hits_nchan, nfolded):
        """hits: [nsub][hits_nchan][nbin] int32; nfolded: one int64 on the device."""
        self.ctx.check(lib().dspx_fold_execute_zeroed(self.h, ptr(x), chan_stride, pol_stride, nchan, npol, ndim,
                                                      ptr(prof), ptr(hits), hits_nchan, ptr(nfolded)))


class Filterbank:
    """dspx_fb: the convolving filterbank engine.  Windows of 1024-4096
    points (sub-bands of 32-512 bins) run as one kernel per window; windows
    of 2^13-2^22 points whose sub-bands hold at least 8*N1 bins (N1 the
    column length of the N1*N2 split) run as three kernels (column
    transform, row transform with the response and the sub-bands' backward
    rows, backward columns with the discard); other lengths, or any length
    with the environment DSPX_FB_FORM=library, run as batches of
    FFT-library transforms with kernels for the response product and the
    overlap discard."""

    def __init__(self, ctx, nchan, npol, nsub, fres, nfilt_pos=0, nfilt_neg=0):
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        h = P()
        ctx.check(lib().dspx_fb_create(ctx.h, C.byref(FbDesc(nchan, npol, nsub, fres, nfilt_pos, nfilt_neg)),
                                       C.byref(h)))
        self.h = h
        _register(self, False)
        self.nchan, self.npol, self.nsub, self.fres = nchan, npol, nsub, fres

    def close(self):
        if self.h:
            self._lib.dspx_fb_destroy(self.h)
            self.h = None

    __del__ = close

    def set_response(self, resp):
        import numpy as np
        r = np.ascontiguousarray(resp, np.complex64)
        assert r.size == self.nchan * self.nsub * self.fres
        self.ctx.check(lib().dspx_fb_set_response(self.h, r.ctypes.data))

    def geometry(self, ndat):
        npart, nout = U64(), U64()
        self.ctx.check(lib().dspx_fb_geometry(self.h, ndat, C.byref(npart), C.byref(nout)))
        return npart.value, nout.value

    def perform(self, x, ics, ips, ndat, out, ocs, ops, zdm=None):
        """zdm: the zero-DM output (same layout as out): each window's
        sub-bands transformed back without the response."""
        if zdm is None:
            self.ctx.check(lib().dspx_fb_perform(self.h, ptr(x), ics, ips, ndat, ptr(out), ocs, ops))
        else:
            self.ctx.check(lib().dspx_fb_perform_zdm(self.h, ptr(x), ics, ips, ndat, ptr(out), ocs, ops,
                                                     ptr(zdm)))

    def batches(self):
        """Library-form batches of the last perform (0: the one-kernel form)."""
        n = U32()
        _check(lib().dspx_fb_batches(self.h, C.byref(n)))
        return n.value


class SpectralKurtosis:
    """dspx_sk: one resolution of SpectralKurtosis (compute / detect / mask).
    Tensors are device buffers: sums [npart][nchan][npol][2] float32, counts
    [npart][nchan] int32 (read as uint32), mask [npart][nchan] uint8."""

    def __init__(self, ctx, nchan, npol, M, Nd=1, omit_outliers=False, noverlap=1):
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        self.nchan, self.npol, self.M, self.noverlap = nchan, npol, M, noverlap
        h = P()
        ctx.check(lib().dspx_sk_create(ctx.h, C.byref(SkDesc(nchan, npol, M, Nd, int(omit_outliers), noverlap)),
                                       C.byref(h)))
        self.h = h
        _register(self, False)

    def close(self):
        if self.h:
            self._lib.dspx_sk_destroy(self.h)
            self.h = None

    __del__ = close

    def set_limits(self, lower, upper, symmetric, pcsd, std_devs=None, Nd=1):
        """Thresholds by block count 0..M.  With std_devs, detect() also
        evaluates SKLimits(total) for the channel totals above M that
        detect_tscr meets (get_limits, SpectralKurtosis.C:989-1012, 1045)."""
        import numpy as np
        lo, up, sy = (np.ascontiguousarray(a, np.float32) for a in (lower, upper, symmetric))
        assert lo.size == up.size == sy.size == self.M + 1
        self.ctx.check(lib().dspx_sk_set_limits(self.h, lo.ctypes.data, up.ctypes.data, sy.ctypes.data, pcsd))
        self.std_devs, self.Nd = std_devs, Nd
        self._tscr = {}

    def blocks(self, ndat):
        """(npart, ndat_out) of a series of ndat samples."""
        npart, nout = U64(), U64()
        _check(lib().dspx_sk_blocks(self.h, ndat, C.byref(npart), C.byref(nout)))
        return npart.value, nout.value

    def set_channels(self, include=None):
        """include: per-channel flags (the averaging detectors' channels), or
        None for every channel."""
        import numpy as np
        if include is None:
            self.ctx.check(lib().dspx_sk_set_channels(self.h, None))
            return
        inc = np.ascontiguousarray(include, np.uint8)
        assert inc.size == self.nchan
        self.ctx.check(lib().dspx_sk_set_channels(self.h, inc.ctypes.data))

    def tscr_totals(self, counts, npart):
        import numpy as np
        t = np.zeros(self.nchan, np.uint64)
        self.ctx.check(lib().dspx_sk_tscr_totals(self.h, ptr(counts), npart, t.ctypes.data))
        return t

    def set_tscr_limits(self, limits):
        """limits: {total: (lower, upper)} for totals above M."""
        import numpy as np
        tot = np.array(sorted(limits), np.uint64)
        lo = np.array([limits[int(t)][0] for t in tot], np.float32)
        up = np.array([limits[int(t)][1] for t in tot], np.float32)
        self.ctx.check(lib().dspx_sk_set_tscr_limits(self.h, tot.size, tot.ctypes.data, lo.ctypes.data,
                                                     up.ctypes.data))

    def compute(self, x, chan_stride, pol_stride, npart, sums, counts, zeroed=False):
        self.ctx.check(lib().dspx_sk_compute(self.h, ptr(x), chan_stride, pol_stride, npart, int(zeroed),
                                             ptr(sums), ptr(counts)))

    def compute_raw(self, raw, nbit, t0, npart, sums, counts, scale=1.0, offset_binary=False):
        self.ctx.check(lib().dspx_sk_compute_raw(self.h, ptr(raw), nbit, int(offset_binary), scale, t0, npart,
                                                 ptr(sums), ptr(counts)))

    def detect(self, sums, counts, npart, flags, mask):
        if not (flags & 4) or getattr(self, "std_devs", None) is None:
            self.ctx.check(lib().dspx_sk_detect(self.h, ptr(sums), ptr(counts), npart, flags, ptr(mask)))
            return
        # detect_skfb / detect_fscr, then the totals they leave, SKLimits of
        # the totals not seen before, then detect_tscr
        self.ctx.check(lib().dspx_sk_detect(self.h, ptr(sums), ptr(counts), npart, flags & 3, ptr(mask)))
        need = sorted({int(t) for t in self.tscr_totals(counts, npart)
                       if self.M < t < 32768 and int(t) not in self._tscr})
        if need:
            from .sklimits import SKLimits
            for t in need:
                L = SKLimits(t, self.std_devs, self.Nd, self.npol)
                self._tscr[t] = (L.lower, L.upper)
            self.set_tscr_limits(self._tscr)
        self.ctx.check(lib().dspx_sk_detect_tscr(self.h, ptr(sums), ptr(counts), npart, ptr(mask)))

    def mask(self, mask, x, ics, ips, out, ocs, ops, npart):
        self.ctx.check(lib().dspx_sk_mask(self.h, ptr(mask), ptr(x), ics, ips, ptr(out), ocs, ops, npart))

    # several resolutions (adaptive SK): the engine's M is the first
    def add_resolution(self, M, noverlap, lower, upper, symmetric, include=None, std_devs=None):
        """A coarser resolution (M a multiple of the previous one's, block
        offset M/noverlap a multiple of the previous one's) with its
        thresholds by count 0..M; std_devs: the level detect_res evaluates
        SKLimits at for the totals above M its time-scrunched test meets (the
        last resolution's)."""
        import numpy as np
        lo, up, sy = (np.ascontiguousarray(a, np.float32) for a in (lower, upper, symmetric))
        assert lo.size == up.size == sy.size == M + 1
        inc = None if include is None else np.ascontiguousarray(include, np.uint8)
        self.ctx.check(lib().dspx_sk_add_resolution(self.h, M, noverlap, lo.ctypes.data, up.ctypes.data,
                                                    sy.ctypes.data, None if inc is None else inc.ctypes.data))
        self._res = getattr(self, "_res", []) + [(M, std_devs)]
        self._tscr = {}

    def clear_resolutions(self):
        self.ctx.check(lib().dspx_sk_clear_resolutions(self.h))
        self._res = []
        self._tscr = {}

    def blocks_res(self, ndat):
        """(npart of the first resolution, ndat_out of the masked output)."""
        npart, nout = U64(), U64()
        _check(lib().dspx_sk_blocks_res(self.h, ndat, C.byref(npart), C.byref(nout)))
        return npart.value, nout.value

    def detect_res(self, sums, counts, ndat, flags, mask):
        """Every resolution's detectors over the first resolution's sums of a
        series of ndat samples; the tscr thresholds of totals above the last
        M evaluated with SKLimits where needed."""
        res = getattr(self, "_res", [])
        sd = res[-1][1] if res else getattr(self, "std_devs", None)
        Mlast = res[-1][0] if res else self.M
        if not (flags & 4) or sd is None:
            self.ctx.check(lib().dspx_sk_detect_res(self.h, ptr(sums), ptr(counts), ndat, flags, ptr(mask)))
            return
        import numpy as np
        self.ctx.check(lib().dspx_sk_detect_res(self.h, ptr(sums), ptr(counts), ndat, flags & 3, ptr(mask)))
        t = np.zeros(self.nchan, np.uint64)
        self.ctx.check(lib().dspx_sk_res_totals(self.h, t.ctypes.data))
        need = sorted({int(x) for x in t if Mlast < x < 32768 and int(x) not in self._tscr})
        if need:
            from .sklimits import SKLimits
            for x in need:
                L = SKLimits(x, sd, self.Nd, self.npol)
                self._tscr[x] = (L.lower, L.upper)
            self.set_tscr_limits(self._tscr)
        self.ctx.check(lib().dspx_sk_res_tscr(self.h, ptr(mask)))

    def mask_res(self, mask, x, ics, ips, out, ocs, ops, ndat):
        self.ctx.check(lib().dspx_sk_mask_res(self.h, ptr(mask), ptr(x), ics, ips, ptr(out), ocs, ops, ndat))


class LoadToFold:
    """Unpack -> Convolution -> Detection -> Fold on one block, fused or plain."""

    def __init__(self, ctx, nbit, nchan, npol, ndat, centre_frequency, bandwidth, dm, nbin,
                 state=INTENSITY, ndim=1, chirp_mode=CHIRP_PARITY, form=FORM_FUSED, nfft=0,
                 scale=1.0, offset_binary=False, geom=None):
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        g = geom or geometry(centre_frequency, bandwidth, nchan, dm, nfft)
        self.geom = g
        d = LtfDesc(nbit, int(offset_binary), scale, nchan, npol, 0, ndat,
                    _dedisp(centre_frequency, bandwidth, dm, nchan), g["nfft"], g["impulse_pos"],
                    g["impulse_neg"], chirp_mode, state, ndim, nbin, form)
        h = P()
        ctx.check(lib().dspx_ltf_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        _register(self, False)
        npart, nout = U64(), U64()
        lib().dspx_ltf_geometry(h, C.byref(npart), C.byref(nout))
        self.npart, self.ndat_out = npart.value, nout.value
        self.nbin, self.state, self.ndim, self.nchan, self.npol = nbin, state, ndim, nchan, npol
        self.nsub = 1

    def close(self):
        if self.h:
            self._lib.dspx_ltf_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def nval(self):
        return {INTENSITY: 1, PPQQ: self.npol}.get(self.state, 4)

    def set_fold(self, pps, chunks):
        arr = (FoldChunk * len(chunks))(*[FoldChunk(c[0], c[1], c[2], c[3], 0) for c in chunks])
        self.ctx.check(lib().dspx_ltf_set_fold(self.h, pps, arr, len(chunks)))
        self.nsub = max(c[3] for c in chunks) + 1

    def execute(self, raw, prof, hits=None):
        self.ctx.check(lib().dspx_ltf_execute(self.h, ptr(raw), ptr(prof), ptr(hits)))

    def execute_masked(self, raw, mask, M, prof, hits=None):
        """The SpectralKurtosis-excised fold: mask [nblk][nchan] uint8 over
        the output samples in blocks of M; hits [nsub][nchan][nbin] counts
        the samples whose first detected value is non-zero."""
        self.ctx.check(lib().dspx_ltf_execute_masked(self.h, ptr(raw), ptr(mask), M, int(mask.shape[0]), ptr(prof),
                                                     ptr(hits)))

    def filterbank(self, raw, out):
        self.ctx.check(lib().dspx_ltf_filterbank(self.h, ptr(raw), ptr(out)))

    def stream_reset(self):
        self.ctx.check(lib().dspx_ltf_stream_reset(self.h))

    def stream_next(self):
        """(npart, ndat_out, t_first) of the next stream call; sets
        self.npart / self.ndat_out for it."""
        npart, nout, t = U64(), U64(), U64()
        self.ctx.check(lib().dspx_ltf_stream_next(self.h, C.byref(npart), C.byref(nout), C.byref(t)))
        self.npart, self.ndat_out = npart.value, nout.value
        return npart.value, nout.value, t.value

    def stream_execute(self, raw, prof, hits=None):
        self.ctx.check(lib().dspx_ltf_stream_execute(self.h, ptr(raw), ptr(prof), ptr(hits)))

    def stream_filterbank(self, raw, out):
        self.ctx.check(lib().dspx_ltf_stream_filterbank(self.h, ptr(raw), ptr(out)))

    def set_timing(self, on=True):
        self.ctx.check(lib().dspx_ltf_set_timing(self.h, int(on)))

    def stage_times(self):
        ms = (C.c_float * 8)()
        names = (C.c_char_p * 8)()
        n = I32()
        self.ctx.check(lib().dspx_ltf_stage_times(self.h, ms, names, 8, C.byref(n)))
        return [(names[i].decode(), ms[i]) for i in range(n.value)]


def fold_phi(t_seconds, period, reference_phase=0.0):
    """Fold::get_phi for a constant folding period (Fold.C:975-982), with the
    reference epoch at t = 0."""
    import math
    return math.fmod(t_seconds, period) / period - reference_phase


def fold_chunks(ndat_out, rate, period, nfilt_pos, nsub=1, t0=0.0, reference_phase=0.0):
    """Sub-integration chunks of one block of convolution output: output
    sample i is input sample nfilt_pos + i (Convolution.C:611); the block is
    divided into nsub equal spans (boundaries rounded with rint, as
    TimeDivide does, TimeDivide.C:193-304).  Returns [(idat_start, ndat, phi,
    isub)] with phi of the midpoint of each chunk's first sample."""
    chunks = []
    # rint: round half to even, as Python's round does
    bounds = [int(round(ndat_out * k / nsub)) for k in range(nsub + 1)]
    for k in range(nsub):
        i0, i1 = bounds[k], bounds[k + 1]
        t = t0 + (nfilt_pos + i0 + 0.5) / rate
        chunks.append((i0, i1 - i0, fold_phi(t, period, reference_phase), k))
    return chunks


BATCH_AUTOTUNE = 1


def ltf_desc(nbit, nchan, npol, ndat, centre_frequency, bandwidth, dm, nbin, state=INTENSITY, ndim=1,
             chirp_mode=CHIRP_PARITY, form=FORM_FUSED, nfft=0, nfilt_pos=0, nfilt_neg=0, scale=1.0,
             offset_binary=False):
    """A dspx_ltf_desc (nfft = 0 leaves the length to the batch)."""
    return LtfDesc(nbit, int(offset_binary), scale, nchan, npol, 0, ndat,
                   _dedisp(centre_frequency, bandwidth, dm, nchan, False), nfft, nfilt_pos, nfilt_neg,
                   chirp_mode, state, ndim, nbin, form)


class Batch:
    """dspx_batch: beams of uneven shape through the fused path, executed in
    order or replayed from a captured HIP graph."""

    def __init__(self, ctx, descs, autotune=False, branches=1):
        """branches: streams the beams are split over (1: all in order on
        the context's stream)."""
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        arr = (LtfDesc * len(descs))(*descs)
        h = P()
        flags = (BATCH_AUTOTUNE if autotune else 0) | (int(branches) << 8)
        ctx.check(lib().dspx_batch_create(ctx.h, arr, len(descs), flags, C.byref(h)))
        self.h = h
        _register(self, False)
        self.nbeam = len(descs)
        self.beams = []
        for i in range(self.nbeam):
            d, npart, nout = LtfDesc(), U64(), U64()
            ctx.check(lib().dspx_batch_beam(h, i, C.byref(d), C.byref(npart), C.byref(nout)))
            self.beams.append({"nfft": d.nfft, "nfilt_pos": d.nfilt_pos, "nfilt_neg": d.nfilt_neg,
                               "npart": npart.value, "ndat_out": nout.value, "form": d.form,
                               "nval": 1 if d.state == INTENSITY else (d.npol if d.state == PPQQ else 4)})

    def close(self):
        if self.h:
            self._lib.dspx_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def set_fold(self, i, pps, chunks):
        arr = (FoldChunk * len(chunks))(*[FoldChunk(c[0], c[1], c[2], c[3], 0) for c in chunks])
        self.ctx.check(lib().dspx_batch_set_fold(self.h, i, pps, arr, len(chunks)))

    def _ptrs(self, raws, profs, hits):
        self._keep = (raws, profs, hits)
        r = (P * self.nbeam)(*[ptr(x) for x in raws])
        p = (P * self.nbeam)(*[ptr(x) for x in profs])
        h = (P * self.nbeam)(*[ptr(x) for x in hits]) if hits is not None else None
        return r, p, h

    def execute(self, raws, profs, hits=None):
        r, p, h = self._ptrs(raws, profs, hits)
        self.ctx.check(lib().dspx_batch_execute(self.h, r, p, h))

    def capture(self, raws, profs, hits=None):
        r, p, h = self._ptrs(raws, profs, hits)
        self.ctx.check(lib().dspx_batch_capture(self.h, r, p, h))

    def replay(self):
        self.ctx.check(lib().dspx_batch_replay(self.h))


def rfi_features(ctx, sums, counts, nblk, nchan, npol, out, Nd=1):
    """Features [nblk][4][nchan] (device tensor out) from SK sums and counts."""
    import torch
    scratch = torch.empty(3 * nblk * nchan + nchan, dtype=torch.float32, device="cuda")
    ctx.check(lib().dspx_rfi_features(ctx.h, ptr(sums), ptr(counts), nblk, nchan, npol, Nd, ptr(scratch), ptr(out)))
    return out


class RfiNet:
    """dspx_rfinet: the RFI proposal network (forward, backward, Adam)."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.h = None
        self._lib = lib()
        h = P()
        ctx.check(lib().dspx_rfinet_create(ctx.h, seed, C.byref(h)))
        self.h = h
        _register(self, False)
        self.nparam = int(lib().dspx_rfinet_nparam())
        self.t = 0

    def close(self):
        if self.h:
            self._lib.dspx_rfinet_destroy(self.h)
            self.h = None

    __del__ = close

    def params(self, rounded=False):
        import numpy as np
        out = np.zeros(self.nparam, np.float32)
        self.ctx.check(lib().dspx_rfinet_get_params(self.h, out.ctypes.data, int(rounded)))
        return out

    def set_params(self, w):
        import numpy as np
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nparam
        self.ctx.check(lib().dspx_rfinet_set_params(self.h, w.ctypes.data))

    def forward(self, feat, nblk, nchan, logits):
        self.ctx.check(lib().dspx_rfinet_forward(self.h, ptr(feat), nblk, nchan, ptr(logits)))

    def backward(self, feat, labels, nblk, nchan, pos_weight=1.0, loss=None):
        self.ctx.check(lib().dspx_rfinet_backward(self.h, ptr(feat), ptr(labels), nblk, nchan, pos_weight,
                                                  ptr(loss) if loss is not None else None))

    def grad(self):
        import numpy as np
        out = np.zeros(self.nparam, np.float32)
        self.ctx.check(lib().dspx_rfinet_grad(self.h, out.ctypes.data))
        return out

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.t += 1
        self.ctx.check(lib().dspx_rfinet_adam(self.h, lr, beta1, beta2, eps, self.t))

    def propose(self, logits, n, mask, threshold=0.0):
        self.ctx.check(lib().dspx_rfinet_propose(self.ctx.h, ptr(logits), n, threshold, ptr(mask)))
