"""Throughput of the coherent-dedispersion + fold hot path on one MI355X.

Default workload (the configuration the metric is quoted on): the uneven
batch, 48 beams of 32-512 channels, 2^18-2^22 samples per channel, int8 and
int16, DM 5-1500 at 300-1500 MHz, full Stokes, 512-2048 bins.  One step is
one graph replay of the whole batch (unpack -> overlap-save convolution ->
detection -> fold for every beam), raw input resident in HBM.  The other
workloads (--workload multichan|single, and the engines beside the path)
are kept for comparison.  Prints one JSON line (see the keys below).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload uneven|multichan|single|...]
  python bench.py --copy-probe MIB   # a plain device copy of MIB MiB, for
                                     # calibrating the HBM byte counters

roofline.traffic is the HBM bytes per launch (per step, for the batch) read
by FETCH_SIZE / WRITE_SIZE counter runs (scaled by the same counters'
reading of a copy kernel of known size); it is taken from
profiles/traffic_<workload>.json when that file exists and is null
otherwise.
"""
import argparse
import importlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s dedispersed+folded; % HBM roofline; fused-MFMA vs plain FP32"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # 128 channels x 2 pols x 2^20 complex int8, 400 MHz at 1400 MHz, DM 300,
    # P = 253 ms, 2% duty at 0.1 sigma, full Stokes, 1024 bins x 8 subints
    "multichan": dict(nchan=128, npol=2, ndat=1 << 20, cfreq=1400.0, bw=400.0, dm=300.0, period=0.253,
                      duty=0.02, peak=0.1, sigma=10.0, nbit=8, state=3, ndim=4, nbin=1024, nsub=8),
    # 1 channel x 2 pols x 2^24 complex int8, 16 MHz at 1405 MHz, DM 56.77,
    # P = 89.33 ms, 3% duty at 0.2 sigma, total intensity, 1024 bins
    "single": dict(nchan=1, npol=2, ndat=1 << 24, cfreq=1405.0, bw=16.0, dm=56.77, period=0.08933,
                   duty=0.03, peak=0.2, sigma=10.0, nbit=8, state=0, ndim=1, nbin=1024, nsub=1),
}


def algorithmic_bytes(w, geom, npart):
    """Per-stage algorithmic HBM bytes of one block: stage A reads the raw
    samples of every FFT window (B_in = 2*nbit/8 per sample) and writes 8 B;
    stage B reads and writes 8 B; stage C reads 8 B (the chirp is generated on
    chip; profiles are O(nbin))."""
    n = w["nchan"] * w["npol"] * npart * geom["nfft"]
    b_in = 2 * w["nbit"] / 8
    return {"stageA_unpack_colfft": n * (b_in + 8), "stageB_rowfft_chirp": n * 16,
            "stageC_colifft_detect_fold": n * 8}


def time_path(dspx, ctx, w, geom, raw, form, steps, warmup, timing_steps=3):
    """One step = the block's fold plan (phases -> bin plan, the per-block
    set-up of Fold::transformation) and the block through the whole path."""
    import torch
    L = dspx.LoadToFold(ctx, w["nbit"], w["nchan"], w["npol"], w["ndat"], w["cfreq"], w["bw"], w["dm"],
                        w["nbin"], state=w["state"], ndim=w["ndim"], form=form, geom=geom)
    rate = abs(w["bw"]) / w["nchan"] * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, w["period"], geom["impulse_pos"], nsub=w["nsub"])
    pps = (1.0 / rate) / w["period"]
    nval = L.nval
    prof = torch.zeros((w["nsub"], w["nchan"], nval * w["nbin"]), dtype=torch.float32, device="cuda")
    hits = torch.zeros((w["nsub"], w["nbin"]), dtype=torch.int32, device="cuda")

    def step():
        L.set_fold(pps, chunks)
        L.execute(raw, prof, hits)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    # per-kernel durations from HIP events on the library's stream (torch's
    # current stream here), the fold plan's set-up included
    L.set_timing(True)
    stage = {}
    cur = torch.cuda.current_stream()
    for _ in range(timing_steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(cur)
        L.set_fold(pps, chunks)
        e1.record(cur)
        L.execute(raw, prof, hits)
        torch.cuda.synchronize()
        stage.setdefault("fold_plan", []).append(e0.elapsed_time(e1))
        for name, ms in L.stage_times():
            stage.setdefault(name, []).append(ms)
    L.set_timing(False)
    torch.cuda.synchronize()
    prof_h = prof.cpu().numpy()
    hits_h = hits.cpu().numpy()
    npart = L.npart
    L.close()
    return t, {k: sum(v) / len(v) for k, v in stage.items()}, prof_h, hits_h, npart


def cpu_baseline(w, geom, raw_np, budget_s=20.0, threads=16):
    """The oracle chain (unpack, FP32 overlap-save convolution with its own
    FFT, detection, fold) on the host cores, up to ``threads`` threads (16
    by default: a host CPU share per GPU), as the reference's MultiThread pipelines
    share a block's work.  The threads run the oracle's C functions, which
    release the interpreter lock.  With at least as many channels as
    threads, each thread takes whole channels, in rounds until the time
    budget is spent; with fewer, each channel's FFT windows (parts) are
    split into contiguous time ranges, one per thread, each folded from its
    own first sample's phase.  One warm channel alone gives the single-core
    rate."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from helpers import oracle_load_to_fold
    O.lib()
    rate = abs(w["bw"]) / w["nchan"] * 1e6
    step, novl = geom["step"], geom["noverlap"]
    npart = (w["ndat"] - novl) // step
    import importlib as il
    dspx = il.import_module("dspsr-kanyixia_amd")
    pps = (1.0 / rate) / w["period"]

    def piece(c, pa, pb):
        """Parts [pa, pb) of channel c through the oracle chain."""
        sub = raw_np[pa * step:pb * step + novl, c:c + 1].copy()
        chunks = dspx.fold_chunks((pb - pa) * step, rate, w["period"], geom["impulse_pos"],
                                  nsub=w["nsub"] if (pa, pb) == (0, npart) else 1, t0=pa * step / rate)
        oracle_load_to_fold(O, sub, w["nbit"], 1, w["npol"], (pb - pa) * step + novl, w["cfreq"], w["bw"],
                            w["dm"], geom, w["state"], w["ndim"], w["nbin"], pps, chunks,
                            w["nsub"] if (pa, pb) == (0, npart) else 1)

    piece(0, 0, npart)  # warm: FFT tables, first touch
    t0 = time.perf_counter()
    piece(0, 0, npart)
    t_one = time.perf_counter() - t0
    per_chan = w["npol"] * w["ndat"]  # counted as the GPU figure counts them: the whole input block
    threads = max(1, min(threads, os.cpu_count() or 1))
    done = 0
    with ThreadPoolExecutor(threads) as ex:
        if w["nchan"] >= threads:
            mode = "whole channels"
            t0 = time.perf_counter()
            while done < w["nchan"] and time.perf_counter() - t0 < budget_s:
                n = min(threads, w["nchan"] - done)
                list(ex.map(lambda c: piece(c, 0, npart), range(done, done + n)))
                done += n
            t_all = time.perf_counter() - t0
        else:
            mode = "each channel's parts in %d time ranges" % min(threads, npart)
            nr = min(threads, npart)
            bounds = [npart * k // nr for k in range(nr + 1)]
            t0 = time.perf_counter()
            while done < w["nchan"] and time.perf_counter() - t0 < budget_s:
                list(ex.map(lambda k: piece(done, bounds[k], bounds[k + 1]), range(nr)))
                done += 1
            t_all = time.perf_counter() - t0
    return {"value": done * per_chan / t_all / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "single_core_value": per_chan / t_one / 1e6,
            "sample": "%d of %d channels of the same block (%d pol x %d samples, %d FFT windows each), %s on %d "
                      "threads, oracle chain unpack+convolution+detection+fold, %.1f s wall; one warm channel "
                      "alone %.2f s" % (done, w["nchan"], w["npol"], w["ndat"], npart, mode, threads, t_all,
                                        t_one)}


def draw_uneven(dspx, nbeam, seed=3, ndat_max_log2=22.0):
    """The uneven-batch workload: beams with nchan from {32, 64, 128, 256,
    512}, DM log-uniform 5-1500, centre frequency uniform 300-1500 MHz, band
    min(0.25*fc, 400) MHz, 2^U(18, 22) samples per channel (multiple of 64),
    8- or 16-bit, period log-uniform 5-500 ms, duty 1-5%, 512/1024/2048 bins.
    Draws whose overlap exceeds a quarter of the block, whose reference FFT
    length exceeds the block (the reference needs one FFT per block,
    Convolution.C:485-487), or whose smearing the reference rule rejects, are
    drawn again."""
    import numpy as np
    rng = np.random.default_rng(seed)
    beams = []
    while len(beams) < nbeam:
        nchan = int(rng.choice([32, 64, 128, 256, 512]))
        dm = float(np.exp(rng.uniform(np.log(5.0), np.log(1500.0))))
        fc = float(rng.uniform(300.0, 1500.0))
        bw = min(0.25 * fc, 400.0)
        ndat = int(2 ** rng.uniform(18.0, ndat_max_log2)) // 64 * 64
        nbit = int(rng.choice([8, 16]))
        period = float(np.exp(rng.uniform(np.log(0.005), np.log(0.5))))
        duty = float(rng.uniform(0.01, 0.05))
        nbin = int(rng.choice([512, 1024, 2048]))
        try:
            g = dspx.geometry(fc, bw, nchan, dm)
        except dspx.DspxError:
            continue
        if g["noverlap"] > ndat / 4 or g["nfft"] > ndat:
            continue
        beams.append(dict(nchan=nchan, npol=2, ndat=ndat, cfreq=fc, bw=bw, dm=dm, nbit=nbit, period=period,
                          duty=duty, nbin=nbin, geom=g))
    return beams


def reference_bytes(b, g):
    """Algorithmic HBM bytes of one beam on the reference-rule basis: FFT
    length N, overlap and r = N/step from the reference's geometry rule;
    B_in = 2*nbit/8 per complex sample; A = r*B_in per input sample where the
    whole transform stays on chip (N <= 2^14), r*(B_in + 32) for the
    two-level form (the window written and read back twice)."""
    n = b["nchan"] * b["npol"] * b["ndat"]
    r = g["nfft"] / g["step"]
    b_in = 2 * b["nbit"] / 8
    a = r * b_in if g["nfft"] <= (1 << 14) else r * (b_in + 32)
    return n * a


def cpu_baseline_uneven(beams, raws, dspx, budget_s=20.0, threads=16):
    """The oracle chain (unpack, FP32 overlap-save convolution at the
    reference FFT length with the oracle's own FFT, full-Stokes detection,
    fold) on the host cores, whole channels of the batch's beams taken
    round-robin over the beams (so that the sample follows the batch's mix
    of lengths, channel counts and sample widths), on up to ``threads``
    threads (16 by default: the host CPU share per GPU), until the time
    budget is spent.  The threads run the oracle's C functions, which release
    the interpreter lock."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from helpers import oracle_load_to_fold
    O.lib()
    threads = max(1, min(threads, os.cpu_count() or 1))
    order = []
    for c in range(max(b["nchan"] for b in beams)):
        for i, b in enumerate(beams):
            if c < b["nchan"]:
                order.append((i, c))

    def task(ic):
        i, c = ic
        b, g = beams[i], beams[i]["geom"]
        raw = np.ascontiguousarray(raws[i][:, c:c + 1].cpu().numpy())
        rate = b["bw"] / b["nchan"] * 1e6
        npart = (b["ndat"] - g["noverlap"]) // g["step"]
        chunks = dspx.fold_chunks(npart * g["step"], rate, b["period"], g["impulse_pos"])
        t0 = time.perf_counter()
        # channel c alone: a one-channel band of its centre and width has its chirp
        cf = b["cfreq"] - 0.5 * b["bw"] + (c + 0.5) * b["bw"] / b["nchan"]
        oracle_load_to_fold(O, raw, b["nbit"], 1, b["npol"], b["ndat"], cf, b["bw"] / b["nchan"], b["dm"],
                            g, dspx.STOKES, 4, b["nbin"], (1.0 / rate) / b["period"], chunks, 1)
        return b["npol"] * b["ndat"], time.perf_counter() - t0

    # one warm channel alone: the single-core rate
    n1, t1 = task(order[0])
    done_samp, k = 0, 0
    t_start = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while k < len(order) and time.perf_counter() - t_start < budget_s:
            res = list(ex.map(task, order[k:k + threads]))
            done_samp += sum(r[0] for r in res)
            k += len(res)
    t_all = time.perf_counter() - t_start
    nbeams = len(set(i for i, _ in order[:k]))
    return {"value": round(done_samp / t_all / 1e6, 2), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "single_core_value": round(n1 / t1 / 1e6, 2), "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": "%d channels (%d of the %d beams, taken round-robin over the beams), %.3g complex samples, "
                      "each channel through the oracle chain (unpack, FP32 overlap-save convolution at the "
                      "reference FFT length, Stokes detection, fold) on %d threads, %.1f s wall" % (
                          k, nbeams, len(beams), done_samp, threads, t_all)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def run_uneven(args, dspx, synth, ctx):
    """The uneven batch (BASELINE's headline configuration): the fused batch
    (autotuned FFT lengths, one-pass or three-stage form per beam, replayed
    from a captured graph) against the plain FFT-library form run beam after
    beam, on the same synthetic beams, with the oracle chain on the host's
    cores beside them.  Full Stokes, one sub-integration per beam."""
    import numpy as np
    import torch
    beams = draw_uneven(dspx, args.nbeam, ndat_max_log2=args.uneven_log2)
    nsamp = sum(b["nchan"] * b["npol"] * b["ndat"] for b in beams)
    print("[bench] %d beams, %.3g samples, raw %.1f GB" % (
        len(beams), nsamp, sum(b["nchan"] * b["npol"] * b["ndat"] * 2 * b["nbit"] // 8 for b in beams) / 1e9),
        file=sys.stderr, flush=True)
    raws = []
    for i, b in enumerate(beams):
        raws.append(synth.make_baseband_fast_torch(b["nchan"], b["npol"], b["ndat"], b["cfreq"], b["bw"], b["dm"],
                                                   b["period"], b["duty"], 0.2,
                                                   sigma=10.0 if b["nbit"] == 8 else 300.0, nbit=b["nbit"],
                                                   seed=100 + i))
    torch.cuda.synchronize()
    torch.cuda.empty_cache()  # the generator's temporaries: the batch allocates its scratch with hipMalloc
    print("[bench] input generated", file=sys.stderr, flush=True)
    descs = [dspx.ltf_desc(b["nbit"], b["nchan"], b["npol"], b["ndat"], b["cfreq"], b["bw"], b["dm"], b["nbin"],
                           state=dspx.STOKES, ndim=4) for b in beams]
    t0 = time.perf_counter()
    B = dspx.Batch(ctx, descs, autotune=True, branches=args.branches)
    tune_s = time.perf_counter() - t0
    profs, hits = [], []
    for i, b in enumerate(beams):
        rate = abs(b["bw"]) / b["nchan"] * 1e6
        info = B.beams[i]
        B.set_fold(i, (1.0 / rate) / b["period"], dspx.fold_chunks(info["ndat_out"], rate, b["period"],
                                                                  info["nfilt_pos"]))
        profs.append(torch.zeros((1, b["nchan"], 4 * b["nbin"]), dtype=torch.float32, device="cuda"))
        hits.append(torch.zeros((1, b["nbin"]), dtype=torch.int32, device="cuda"))
    B.execute(raws, profs, hits)
    free, total = torch.cuda.mem_get_info()
    print("[bench] batch ready: %d streams, device memory %.1f of %.1f GB in use" % (
        args.branches, (total - free) / 1e9, total / 1e9), file=sys.stderr, flush=True)
    B.capture(raws, profs, hits)
    for _ in range(args.warmup):
        B.replay()
    stream = torch.cuda.current_stream()  # the library's stream (the context was given it)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        B.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    t_f = (time.perf_counter() - t0) / args.steps
    t_ev = e0.elapsed_time(e1) * 1e-3 / args.steps
    print("[bench] fused batch %.3f ms/step (events %.3f ms; autotune %.1f s)" % (t_f * 1e3, t_ev * 1e3, tune_s),
          file=sys.stderr, flush=True)
    ref_bytes = sum(reference_bytes(b, b["geom"]) for b in beams)
    run_bytes = 0  # the windows as run: the chosen lengths and forms
    forms = {"onepass": 0, "twolevel": 0}
    for i, b in enumerate(beams):
        info = B.beams[i]
        n = b["nchan"] * b["npol"] * info["npart"] * info["nfft"]
        b_in = 2 * b["nbit"] / 8
        one = info["form"] == dspx.FORM_ONEPASS
        run_bytes += n * (b_in if one else b_in + 32)
        forms["onepass" if one else "twolevel"] += 1
    nfft_tuned = [(B.beams[i]["nfft"], b["geom"]["nfft"]) for i, b in enumerate(beams)]
    tuned = [(B.beams[i]["nfft"], B.beams[i]["nfilt_pos"], B.beams[i]["nfilt_neg"]) for i in range(len(beams))]
    B.close()
    t_m = None
    if not args.no_mfma:
        # the same lengths with stage B's 4096-point rows on the matrix cores
        os.environ["DSPX_STAGEB"] = "mfma"
        try:
            d2 = [dspx.ltf_desc(b["nbit"], b["nchan"], b["npol"], b["ndat"], b["cfreq"], b["bw"], b["dm"], b["nbin"],
                                state=dspx.STOKES, ndim=4, nfft=tuned[i][0], nfilt_pos=tuned[i][1],
                                nfilt_neg=tuned[i][2]) for i, b in enumerate(beams)]
            B2 = dspx.Batch(ctx, d2, branches=args.branches)
            for i, b in enumerate(beams):
                rate = abs(b["bw"]) / b["nchan"] * 1e6
                info = B2.beams[i]
                B2.set_fold(i, (1.0 / rate) / b["period"], dspx.fold_chunks(info["ndat_out"], rate, b["period"],
                                                                           info["nfilt_pos"]))
            B2.execute(raws, profs, hits)
            B2.capture(raws, profs, hits)
            B2.replay()
            torch.cuda.synchronize()
            e0.record(stream)
            for _ in range(args.steps):
                B2.replay()
            e1.record(stream)
            torch.cuda.synchronize()
            t_m = e0.elapsed_time(e1) * 1e-3 / args.steps
            B2.close()
        finally:
            del os.environ["DSPX_STAGEB"]
        print("[bench] fused batch, stage B on the matrix cores %.3f ms/step" % (t_m * 1e3), file=sys.stderr,
              flush=True)
    del profs, hits
    torch.cuda.empty_cache()
    t_p = None
    if not args.no_plain:
        t_p = 0.0
        for i, b in enumerate(beams):
            L = dspx.LoadToFold(ctx, b["nbit"], b["nchan"], b["npol"], b["ndat"], b["cfreq"], b["bw"], b["dm"],
                                b["nbin"], state=dspx.STOKES, ndim=4, form=dspx.FORM_PLAIN, geom=b["geom"])
            rate = abs(b["bw"]) / b["nchan"] * 1e6
            L.set_fold((1.0 / rate) / b["period"], dspx.fold_chunks(L.ndat_out, rate, b["period"],
                                                                   b["geom"]["impulse_pos"]))
            prof = torch.zeros((1, b["nchan"], 4 * b["nbin"]), dtype=torch.float32, device="cuda")
            hit = torch.zeros((1, b["nbin"]), dtype=torch.int32, device="cuda")
            L.execute(raws[i], prof, hit)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L.execute(raws[i], prof, hit)
            torch.cuda.synchronize()
            t_p += time.perf_counter() - t0
            L.close()
            del prof, hit
            torch.cuda.empty_cache()
        print("[bench] plain beam by beam %.3f ms" % (t_p * 1e3), file=sys.stderr, flush=True)
    value = nsamp / t_f / 1e6
    traffic = load_traffic("uneven", "batch")
    out = {"metric": METRIC, "value": round(value, 2), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(t_f * 1e3, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (noise plus a pulse dispersed between channels, not within them)",
           "config": {"workload": "uneven batch: %d beams (seed 3), nchan %s, 2^18-2^%g samples per channel, "
                                  "int8/int16, full Stokes, 512-2048 bins, one graph replay per step" % (
                                      len(beams), sorted(set(b["nchan"] for b in beams)), args.uneven_log2),
                      "samples_per_step": nsamp,
                      "form": "fused batch, autotuned nfft, graph replay, beams over %d streams" % args.branches,
                      "forms": forms, "nfft_tuned_vs_reference": nfft_tuned},
           "roofline": {"bound": "hbm", "kernel": "whole batch (graph replay of every beam's kernels)",
                        "achieved": round(ref_bytes / t_ev / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(ref_bytes / t_ev / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "basis": "reference-rule geometry per beam: A = r*B_in (N <= 2^14), r*(B_in + 32) above",
                        "algorithmic_bytes_per_launch": int(ref_bytes), "ms_per_launch_events": round(t_ev * 1e3, 3),
                        "as_run": {"algorithmic_bytes_per_launch": int(run_bytes),
                                   "achieved": round(run_bytes / t_ev / 1e9, 1),
                                   "frac": round(run_bytes / t_ev / 1e9 / HBM_PEAK_GBS, 4),
                                   "basis": "the chosen lengths: raw bytes of every window (one-pass), plus 32 B "
                                            "per window sample (three-stage)"}},
           "autotune_seconds": round(tune_s, 2)}
    if t_p is not None:
        out["plain_fp32"] = {"value": round(nsamp / t_p / 1e6, 2), "ms_per_step": round(t_p * 1e3, 3),
                             "form": "FFT-library form, beam after beam, reference FFT lengths"}
        out["fused_vs_plain"] = round(t_p / t_f, 3)
    if t_m is not None:
        out["fused_mfma"] = {"value": round(nsamp / t_m / 1e6, 2), "ms_per_step": round(t_m * 1e3, 3),
                             "form": "the same batch and lengths with stage B's 4096-point rows as FP16 hi/lo-split "
                                     "matrix-core products (FP32 accumulation); by HIP events"}
        if t_p is not None:
            out["fused_mfma_vs_plain"] = round(t_p / t_m, 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_uneven(beams, raws, dspx, args.cpu_budget)
    print(json.dumps(out))


ENGINE_WORKLOADS = ("sk", "rfinet", "filterbank", "delay")


def _event_ms(fn, reps):
    """Average milliseconds of ``fn`` over ``reps`` calls, by HIP events on
    the current (library) stream, after one untimed call."""
    import torch
    fn()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def run_engine(args, dspx, synth, ctx):
    """The engines beside the dedispersion path, each on device-resident
    synthetic input of the shape it meets in the pipeline, timed by HIP
    events: one JSON line with the throughput and the bound of its main
    kernel (algorithmic bytes, or flops for the network, per call).

      sk:         SpectralKurtosis on 512 channels x 2 pols x 2^20 int8
                  samples with injected RFI, M = 1024: block sums from the
                  raw bytes, the three detectors, the mask of the unpacked
                  float samples (the mask is the HBM-heavy step).
      rfinet:     the proposal network's inference (features, 4 conv layers,
                  threshold) on that data's 1024 x 512 blocks.
      filterbank: the convolving filterbank engine, 128 channels x 2 pols x
                  2^20 float samples into 16 sub-bands of 128 bins.
      delay:      SampleDelay on 128 channels x 2 pols x 2^20 float samples
                  (the -K inter-channel delays of the 128-channel band at
                  DM 100; at DM 300 the band's delay exceeds the block)."""
    import numpy as np
    import torch
    w = args.workload
    reps = max(3, args.steps)
    out = {"metric": None, "value": None, "unit": None, "n_gpus": 1, "steps": reps, "warmup": 1,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "cpu_baseline": None}
    if w in ("sk", "rfinet"):
        nchan, npol, ndat, M = 512, 2, 1 << 20, 1024
        nblk = ndat // M
        raw, truth = synth.make_rfi_baseband_torch(nchan, npol, ndat, M, 11)
        torch.cuda.synchronize()
        S = importlib.import_module("dspsr-kanyixia_amd.sklimits")
        lo, up, sym, pcsd = S.limit_tables(M, 3.0, npol=npol, counts=[M])
        sk = dspx.SpectralKurtosis(ctx, nchan, npol, M, omit_outliers=True)
        sk.set_limits(lo, up, sym, pcsd)
        sums = torch.zeros((nblk, nchan, npol, 2), dtype=torch.float32, device="cuda")
        counts = torch.zeros((nblk, nchan), dtype=torch.int32, device="cuda")
        mask = torch.zeros((nblk, nchan), dtype=torch.uint8, device="cuda")
        nsamp = nchan * npol * ndat
        if w == "sk":
            x = torch.empty((nchan, npol, ndat, 2), dtype=torch.float32, device="cuda")
            x.copy_(raw.permute(1, 2, 0, 3).float())
            y = torch.empty_like(x)
            t_c = _event_ms(lambda: sk.compute_raw(raw, 8, 0, nblk, sums, counts), reps)
            s2, c2 = sums.clone(), counts.clone()

            def det():
                s2.copy_(sums)
                c2.copy_(counts)
                sk.detect(s2, c2, nblk, 7, mask)
            t_d = _event_ms(det, reps)
            t_m = _event_ms(lambda: sk.mask(mask, x, npol * ndat * 2, ndat * 2, y, npol * ndat * 2, ndat * 2, nblk),
                            reps)
            t = t_c + t_d + t_m
            mbytes = 2 * nsamp * 8 + nblk * nchan
            frac_flag = float(mask.float().mean())
            out.update({"metric": "Msamples/s spectral kurtosis (sums from raw, detect, mask)",
                        "value": round(nsamp / (t * 1e-3) / 1e6, 2), "unit": "Msamples/s",
                        "ms_per_step": round(t, 4),
                        "config": {"workload": "sk: 512 chan x 2 pol x 2^20 complex int8 with injected RFI, "
                                               "M 1024, 3 sigma, skfb+fscr+tscr with omit_outliers",
                                   "flagged_fraction": round(frac_flag, 4)},
                        "stages_ms": {"compute_raw": round(t_c, 4), "detect": round(t_d, 4), "mask": round(t_m, 4)},
                        "roofline": {"bound": "hbm", "kernel": "mask", "achieved": round(mbytes / t_m / 1e6, 1),
                                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(mbytes / t_m / 1e6 / HBM_PEAK_GBS, 4), "traffic": None,
                                     "compute_raw_GBps": round(nsamp * 2 / t_c / 1e6, 1)}})
        else:
            sk.compute_raw(raw, 8, 0, nblk, sums, counts)
            feat = torch.zeros((nblk, 4, nchan), dtype=torch.float32, device="cuda")
            logits = torch.zeros((nblk, nchan), dtype=torch.float32, device="cuda")
            net = dspx.RfiNet(ctx, seed=7)

            def infer():
                dspx.rfi_features(ctx, sums, counts, nblk, nchan, npol, feat)
                net.forward(feat, nblk, nchan, logits)
                net.propose(logits, nblk * nchan, mask)
            t = _event_ms(infer, reps)
            t_f = _event_ms(lambda: net.forward(feat, nblk, nchan, logits), reps)
            K, Wd, F = 19, 32, 4
            flops = 2 * nblk * nchan * K * (F * Wd + 2 * Wd * Wd + Wd)
            net.close()
            out.update({"metric": "block-channel positions/s, RFI proposal inference (features, network, threshold)",
                        "value": round(nblk * nchan / (t * 1e-3), 1), "unit": "positions/s",
                        "ms_per_step": round(t, 4),
                        "config": {"workload": "rfinet: SK sums of 512 chan x 2^20 samples, M 1024 -> 1024 x 512 "
                                               "positions, 4 conv layers (4-32-32-32-1, kernel 19 across channels)"},
                        "roofline": {"bound": "mfma", "kernel": "forward (4 conv layers)",
                                     "achieved": round(flops / t_f / 1e9, 2), "peak": 2500.0, "unit": "TFLOP/s",
                                     "frac": round(flops / t_f / 1e9 / 2500.0, 4), "traffic": None,
                                     "note": "algorithmic flops of the 4 layers over the forward's time; the two "
                                             "32-32 layers (94% of them) run as BF16 16x16x32 matrix products with "
                                             "the activations split hi/lo (twice the matrix flops), the first and "
                                             "last layers on the VALU; peak: dense BF16 matrix cores"}})
        sk.close()
    elif w in ("filterbank", "delay"):
        nchan, npol, ndat = 128, 2, 1 << 20
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        x = torch.randn((nchan, npol, ndat, 2), generator=g, device="cuda") * 10
        nsamp = nchan * npol * ndat
        if w == "filterbank":
            nsub, fres = 16, 128
            fb = dspx.Filterbank(ctx, nchan, npol, nsub, fres)
            resp = np.exp(1j * np.random.default_rng(1).uniform(0, 2 * np.pi, (nchan * nsub, fres)))
            fb.set_response(resp.astype(np.complex64))
            npart, nout = fb.geometry(ndat)  # nout samples per output channel
            y = torch.empty((nchan * nsub, npol, nout, 2), dtype=torch.float32, device="cuda")
            run = lambda: fb.perform(x, npol * ndat * 2, ndat * 2, ndat, y, npol * nout * 2, nout * 2)
            t = _event_ms(run, reps)
            # the library form (hipFFT batches, DSPX_FB_FORM=library) timed on the same data
            prev = os.environ.get("DSPX_FB_FORM")
            os.environ["DSPX_FB_FORM"] = "library"
            try:
                t_lib = _event_ms(run, reps)
            finally:
                if prev is None:
                    del os.environ["DSPX_FB_FORM"]
                else:
                    os.environ["DSPX_FB_FORM"] = prev
            fb.close()
            abytes = nsamp * 8 + nchan * nsub * npol * nout * 8
            out.update({"metric": "Msamples/s convolving filterbank (one kernel per window)",
                        "value": round(nsamp / (t * 1e-3) / 1e6, 2), "unit": "Msamples/s", "ms_per_step": round(t, 4),
                        "library_form": {"ms_per_step": round(t_lib, 4),
                                         "value": round(nsamp / (t_lib * 1e-3) / 1e6, 2)},
                        "config": {"workload": "filterbank: 128 chan x 2 pol x 2^20 complex float -> 16 sub-bands "
                                               "x 128 bins per channel", "npart": npart, "ndat_out": nout},
                        "roofline": {"bound": "hbm", "kernel": "k_fb_fused<2048, 128> (float in, float out)",
                                     "achieved": round(abytes / t / 1e6, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(abytes / t / 1e6 / HBM_PEAK_GBS, 4),
                                     "traffic": load_traffic("filterbank", "k_fb_fused")}})
        else:
            sd = dspx.sample_delays(1400.0, 400.0, nchan, 100.0)
            dl = dspx.SampleDelay(ctx, nchan, npol, sd["applied"])
            nret = ndat - int(sd["total"])
            y = torch.empty((nchan, npol, nret, 2), dtype=torch.float32, device="cuda")
            t = _event_ms(lambda: dl.retard(x, npol * ndat * 2, ndat * 2, 2, nret, y, npol * nret * 2, nret * 2), reps)
            dl.close()
            abytes = 2 * nchan * npol * nret * 8
            out.update({"metric": "Msamples/s SampleDelay (inter-channel delays, -K)",
                        "value": round(nchan * npol * nret / (t * 1e-3) / 1e6, 2), "unit": "Msamples/s",
                        "ms_per_step": round(t, 4),
                        "config": {"workload": "delay: 128 chan x 2 pol x 2^20 complex float, 400 MHz at 1400 MHz, "
                                               "DM 100", "total_delay": int(sd["total"])},
                        "roofline": {"bound": "hbm", "kernel": "k_retard", "achieved": round(abytes / t / 1e6, 1),
                                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(abytes / t / 1e6 / HBM_PEAK_GBS, 4), "traffic": None}})
    print(json.dumps(out))


def load_traffic(workload, kernel):
    """HBM bytes per launch of ``kernel`` from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else int(k["hbm_bytes_per_launch"])


def copy_probe(mib, reps=5):
    """A device-to-device copy of ``mib`` MiB: reads and writes exactly that
    many bytes, the yardstick for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="uneven", choices=["uneven"] + list(WORKLOADS) + list(ENGINE_WORKLOADS))
    ap.add_argument("--nbeam", type=int, default=48)
    ap.add_argument("--uneven-log2", type=float, default=22.0, help="largest log2 samples per channel (uneven)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-mfma", action="store_true", help="skip the matrix-core stage B timing (uneven)")
    ap.add_argument("--branches", type=int, default=4,
                    help="streams the uneven batch's beams are split over (4: one per hardware queue; 6 and 8 "
                         "measure the same)")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this path has no multi-GPU form here "
              "(channels and beams shard without exchange)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import numpy as np
    import torch
    dspx = importlib.import_module("dspsr-kanyixia_amd")
    synth = importlib.import_module("dspsr-kanyixia_amd.synth")
    torch.cuda.init()
    # torch's fills and the library's kernels on one stream (the default
    # stream's handle is 0, for which the library would create its own)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx = dspx.Context(0, stream.cuda_stream)
    if args.workload == "uneven":
        run_uneven(args, dspx, synth, ctx)
        ctx.close()
        return
    if args.workload in ENGINE_WORKLOADS:
        run_engine(args, dspx, synth, ctx)
        ctx.close()
        return
    w = WORKLOADS[args.workload]
    geom = dspx.geometry(w["cfreq"], w["bw"], w["nchan"], w["dm"])
    print("[bench] workload %s geometry %s" % (args.workload, geom), file=sys.stderr, flush=True)
    raw = synth.make_baseband_torch(w["nchan"], w["npol"], w["ndat"], w["cfreq"], w["bw"], w["dm"], w["period"],
                                    w["duty"], w["peak"], sigma=w["sigma"], nbit=w["nbit"], seed=1)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()  # the generator's temporaries: the batch allocates its scratch with hipMalloc
    print("[bench] input generated", file=sys.stderr, flush=True)
    nsamp = w["nchan"] * w["npol"] * w["ndat"]

    t_f, st_f, prof_f, hits_f, npart = time_path(dspx, ctx, w, geom, raw, dspx.FORM_FUSED, args.steps, args.warmup)
    print("[bench] fused %.3f ms/step stages %s" % (t_f * 1e3, st_f), file=sys.stderr, flush=True)
    plain = None
    if not args.no_plain:
        sp = max(3, args.steps // 4)
        t_p, st_p, prof_p, hits_p, _ = time_path(dspx, ctx, w, geom, raw, dspx.FORM_PLAIN, sp, 1)
        nf, npl = args.warmup + args.steps + 3, 1 + sp + 3  # executions accumulated in each
        pf, pp = prof_f / nf, prof_p / npl
        plain = {"value": nsamp / t_p / 1e6, "ms_per_step": t_p * 1e3, "stages_ms": st_p,
                 "profile_rel_rms_vs_fused": float(np.sqrt(np.sum((pf - pp) ** 2) / np.sum(pp ** 2))),
                 "hits_equal": bool(np.array_equal(hits_f // nf, hits_p // npl))}
        print("[bench] plain %.3f ms/step stages %s" % (t_p * 1e3, st_p), file=sys.stderr, flush=True)

    value = nsamp / t_f / 1e6
    abytes = algorithmic_bytes(w, geom, npart)
    dom = max(st_f, key=lambda k: st_f[k] if k in abytes else -1)
    ach = abytes[dom] / (st_f[dom] * 1e-3) / 1e9
    total_alg = sum(abytes.values())
    r = geom["nfft"] / geom["step"]
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": load_traffic(args.workload, dom), "kernel": dom,
            "kernel_ms": round(st_f[dom], 4), "algorithmic_bytes_per_launch": int(abytes[dom]),
            "path_achieved": round(total_alg / t_f / 1e9, 1),
            "path_frac": round(total_alg / t_f / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_sample": round(total_alg / nsamp, 3), "r": round(r, 5)}
    out = {"metric": METRIC, "value": round(value, 2), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(t_f * 1e3, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": args.workload + ": %d chan x %d pol x %d complex int%d, %g MHz at %g MHz, DM %g, "
                                                 "state %d ndim %d, %d bins x %d subints" % (
                                                     w["nchan"], w["npol"], w["ndat"], w["nbit"], w["bw"], w["cfreq"],
                                                     w["dm"], w["state"], w["ndim"], w["nbin"], w["nsub"]),
                      "nfft": geom["nfft"], "noverlap": geom["noverlap"], "npart": npart, "form": "fused"},
           "roofline": roof, "stages_ms": {k: round(v, 4) for k, v in st_f.items()}}
    if plain is not None:
        out["plain_fp32"] = {"value": round(plain["value"], 2), "ms_per_step": round(plain["ms_per_step"], 4),
                             "stages_ms": {k: round(v, 4) for k, v in plain["stages_ms"].items()},
                             "profile_rel_rms_vs_fused": plain["profile_rel_rms_vs_fused"],
                             "hits_equal": plain["hits_equal"]}
        out["fused_vs_plain"] = round(value / plain["value"], 3)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, geom, raw.cpu().numpy(), args.cpu_budget)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
