"""Seeded synthetic baseband: complex Gaussian noise quantised to 8 or 16 bits
(TFP [ndat][nchan][npol][2]) with a periodic pulse whose pulsed noise is
dispersed by the cold-plasma transfer function (the conjugate of the
dedispersion chirp, evaluated in FP64 on each channel's full-length FFT).

numpy generates the noise (PCG64 via numpy.random.Generator, seed given);
the dispersion of the pulsed component uses torch on the GPU when one is
present (``device``), numpy FFTs otherwise.  The generator is not part of the
timed work of any benchmark.
"""
import math

import numpy as np

DM_DISPERSION = 2.41e-4  # Dedispersion.C:32


def chirp_phase(freq, chan_freq, bw, dm):
    """Dedispersion phase (rad) at offset freq (MHz) from chan_freq: the
    closed form of Dedispersion::build_compute (Dedispersion.C:166-186)."""
    sign = math.copysign(1.0, bw)
    dpm = 1e6 * dm / DM_DISPERSION
    coeff = -sign * 2 * math.pi * dpm / (chan_freq * chan_freq)
    return coeff * freq * freq / (chan_freq + freq)


def channel_centres(cfreq, bw, nchan):
    base = cfreq - 0.5 * bw + 0.5 * bw / nchan
    return np.array([base + i * bw / nchan for i in range(nchan)])


def pulse_envelope(ndat, rate, period, duty, phase0=0.5, t0=0.0):
    """Gaussian pulse profile (FWHM = duty * period) sampled at the sample
    midpoints (t0 + (i + 0.5) / rate), peak 1 at pulse phase phase0."""
    t = t0 + (np.arange(ndat, dtype=np.float64) + 0.5) / rate
    ph = np.mod(t / period, 1.0)
    d = np.mod(ph - phase0 + 0.5, 1.0) - 0.5
    sig = duty / 2.3548200450309493
    return np.exp(-0.5 * (d / sig) ** 2)


def make_baseband(nchan, npol, ndat, cfreq, bw, dm, period, duty, peak_sigma, sigma=10.0,
                  nbit=8, seed=1, device=None, phase0=0.5):
    """Returns raw TFP bytes as a numpy array of int8 [ndat, nchan, npol, 2]
    (nbit 8) or int16 (nbit 16).  The pulsed component has RMS
    peak_sigma*sigma at the pulse peak before dispersion."""
    rng = np.random.default_rng(seed)
    rate = abs(bw) / nchan * 1e6
    env = pulse_envelope(ndat, rate, period, duty, phase0)
    cen = channel_centres(cfreq, bw, nchan)
    chanwidth = bw / nchan
    lim = 127 if nbit == 8 else 32767
    dt = np.int8 if nbit == 8 else np.int16
    out = np.empty((ndat, nchan, npol, 2), dt)
    # FFT bin k of an ndat-point channel stream sits at offset f_k (MHz)
    k = np.fft.fftfreq(ndat, d=1.0) * ndat
    fk = k * (chanwidth / ndat)
    use_torch = device is not None
    if use_torch:
        import torch
    for ic in range(nchan):
        ph = chirp_phase(fk, cen[ic], bw, dm)
        for ip in range(npol):
            noise = rng.standard_normal((ndat, 2), dtype=np.float32) * sigma
            puls = rng.standard_normal((ndat, 2), dtype=np.float32)
            amp = (peak_sigma * sigma / math.sqrt(2.0)) * env
            if use_torch:
                pc = torch.complex(torch.from_numpy(puls[:, 0]).to(device, torch.float64),
                                   torch.from_numpy(puls[:, 1]).to(device, torch.float64))
                pc = pc * torch.from_numpy(amp).to(device)
                H = torch.polar(torch.ones(ndat, dtype=torch.float64, device=device),
                                torch.from_numpy(-ph).to(device))
                disp = torch.fft.ifft(torch.fft.fft(pc) * H)
                d = torch.stack([disp.real, disp.imag], -1).float().cpu().numpy()
            else:
                pc = (puls[:, 0] + 1j * puls[:, 1]).astype(np.complex128) * amp
                disp = np.fft.ifft(np.fft.fft(pc) * np.exp(-1j * ph))
                d = np.stack([disp.real, disp.imag], -1)
            v = np.rint(noise + d)
            np.clip(v, -lim, lim, out=v)
            out[:, ic, ip, :] = v.astype(dt)
    return out


def make_baseband_torch(nchan, npol, ndat, cfreq, bw, dm, period, duty, peak_sigma, sigma=10.0,
                        nbit=8, seed=1, device="cuda", phase0=0.5):
    """Device-side variant of make_baseband for large workloads: the same
    model, generated with torch's Philox generator on ``device``.  Returns a
    device tensor [ndat, nchan, npol, 2] of int8 / int16."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    rate = abs(bw) / nchan * 1e6
    env = torch.from_numpy(pulse_envelope(ndat, rate, period, duty, phase0)).to(device)
    amp = (peak_sigma * sigma / math.sqrt(2.0)) * env
    cen = channel_centres(cfreq, bw, nchan)
    chanwidth = bw / nchan
    k = torch.fft.fftfreq(ndat, d=1.0, device=device, dtype=torch.float64) * ndat
    fk = k * (chanwidth / ndat)
    lim = 127 if nbit == 8 else 32767
    dt = torch.int8 if nbit == 8 else torch.int16
    out = torch.empty((ndat, nchan, npol, 2), dtype=dt, device=device)
    sign = math.copysign(1.0, bw)
    dpm = 1e6 * dm / DM_DISPERSION
    for ic in range(nchan):
        cf = float(cen[ic])
        coeff = -sign * 2 * math.pi * dpm / (cf * cf)
        ph = coeff * fk * fk / (cf + fk)
        H = torch.polar(torch.ones_like(ph), -ph)
        for ip in range(npol):
            noise = torch.randn((ndat, 2), generator=g, device=device, dtype=torch.float32) * sigma
            puls = torch.randn((ndat, 2), generator=g, device=device, dtype=torch.float64)
            pc = torch.complex(puls[:, 0], puls[:, 1]) * amp
            disp = torch.fft.ifft(torch.fft.fft(pc) * H)
            v = noise + torch.stack([disp.real, disp.imag], -1).float()
            out[:, ic, ip, :] = torch.clamp(torch.round(v), -lim, lim).to(dt)
    return out


def make_baseband_fast_torch(nchan, npol, ndat, cfreq, bw, dm, period, duty, peak_sigma, sigma=10.0,
                             nbit=8, seed=1, device="cuda", phase0=0.5, chunk=1 << 16):
    """Large-workload generator: complex Gaussian noise plus a pulsed noise
    component whose arrival in each channel is delayed by the cold-plasma
    delay at the channel centre (relative to the top of the band), i.e.
    dispersed between channels but not within them.  Generated on ``device``
    in time chunks (torch Philox, seeded); returns [ndat, nchan, npol, 2]
    int8 / int16.  For throughput workloads, where the time a full in-channel
    dispersion would take matters and the detail of the pulse does not."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    rate = abs(bw) / nchan * 1e6
    cen = channel_centres(cfreq, bw, nchan)
    ftop = float(np.max(cen))
    dpm = dm / DM_DISPERSION
    delay = torch.tensor([dpm * (1.0 / (f * f) - 1.0 / (ftop * ftop)) for f in cen], dtype=torch.float64,
                         device=device)
    lim = 127 if nbit == 8 else 32767
    dt = torch.int8 if nbit == 8 else torch.int16
    out = torch.empty((ndat, nchan, npol, 2), dtype=dt, device=device)
    sig = duty / 2.3548200450309493
    amp = peak_sigma * sigma / math.sqrt(2.0)
    for t0 in range(0, ndat, chunk):
        T = min(chunk, ndat - t0)
        t = (torch.arange(t0, t0 + T, device=device, dtype=torch.float64) + 0.5) / rate
        ph = torch.remainder((t[:, None] - delay[None, :]) / period, 1.0)
        d = torch.remainder(ph - phase0 + 0.5, 1.0) - 0.5
        env = (amp * torch.exp(-0.5 * (d / sig) ** 2)).float()  # [T, nchan]
        v = torch.randn((T, nchan, npol, 2), generator=g, device=device) * sigma
        v += torch.randn((T, nchan, npol, 2), generator=g, device=device) * env[:, :, None, None]
        out[t0:t0 + T] = torch.clamp(torch.round(v), -lim, lim).to(dt)
    return out


def make_rfi_baseband_torch(nchan, npol, ndat, M, seed, sigma=10.0, bw=256.0, tone_frac=0.05,
                            impulse_frac=0.01, am_frac=0.01, device="cuda", chunk=1 << 15):
    """Complex Gaussian noise (sigma per component, quantised to int8, TFP
    [ndat][nchan][npol][2]) with injected interference: constant-amplitude
    tones (2-20 sigma) filling tone_frac of the channels, broadband impulses
    (5-50 sigma, 1-64 samples) in impulse_frac of the M-sample blocks, and
    50 Hz amplitude modulation (depth 0.5-1) in am_frac of the channels; the
    channel sample rate is bw/nchan MHz.  Returns (raw, truth) with truth
    [nblk][nchan] bool marking the blocks that hold injected interference."""
    import torch
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    nblk = ndat // M
    rate = bw / nchan * 1e6
    chans = rng.permutation(nchan)
    ntone = max(1, int(round(tone_frac * nchan)))
    nam = max(1, int(round(am_frac * nchan)))
    tone_ch, am_ch = chans[:ntone], chans[ntone:ntone + nam]
    tone_amp = rng.uniform(2, 20, ntone) * sigma
    tone_f = rng.uniform(-0.5, 0.5, ntone)
    tone_ph = rng.uniform(0, 2 * np.pi, (ntone, npol))
    am_depth = rng.uniform(0.5, 1.0, nam)
    am_ph = rng.uniform(0, 2 * np.pi, nam)
    nimp = max(1, int(round(impulse_frac * nblk)))
    imp_blk = np.sort(rng.choice(nblk, nimp, replace=False))
    imp_len = rng.integers(1, 65, nimp)
    imp_amp = rng.uniform(5, 50, nimp) * sigma
    imp_t0 = imp_blk * M + rng.integers(0, np.maximum(M - imp_len, 0) + 1)
    truth = np.zeros((nblk, nchan), bool)
    truth[:, tone_ch] = True
    truth[:, am_ch] = True
    truth[imp_blk, :] = True
    out = torch.empty((ndat, nchan, npol, 2), dtype=torch.int8, device=device)
    tch = torch.as_tensor(tone_ch, device=device)
    ach = torch.as_tensor(am_ch, device=device)
    tamp = torch.as_tensor(tone_amp, dtype=torch.float32, device=device)
    tf = torch.as_tensor(tone_f, dtype=torch.float64, device=device)
    tph = torch.as_tensor(tone_ph, dtype=torch.float64, device=device)
    adep = torch.as_tensor(am_depth, dtype=torch.float32, device=device)
    aph = torch.as_tensor(am_ph, dtype=torch.float64, device=device)
    for t0 in range(0, ndat, chunk):
        T = min(chunk, ndat - t0)
        t = torch.arange(t0, t0 + T, device=device, dtype=torch.float64)
        v = torch.randn((T, nchan, npol, 2), generator=g, device=device) * sigma
        mod = 1 + adep[None, :] * torch.sin(2 * np.pi * 50.0 * t[:, None] / rate + aph[None, :]).float()
        v[:, ach] *= mod[:, :, None, None]
        ph = 2 * np.pi * tf[None, :, None] * t[:, None, None] + tph[None, :, :]  # [T][ntone][npol]
        v[:, tch, :, 0] += tamp[None, :, None] * torch.cos(ph).float()
        v[:, tch, :, 1] += tamp[None, :, None] * torch.sin(ph).float()
        for j in range(nimp):
            a, b = int(imp_t0[j]), int(imp_t0[j] + imp_len[j])
            lo, hi = max(a, t0), min(b, t0 + T)
            if lo < hi:
                v[lo - t0:hi - t0] += torch.randn((hi - lo, nchan, npol, 2), generator=g, device=device) * float(
                    imp_amp[j] / np.sqrt(2.0))
        out[t0:t0 + T] = torch.clamp(torch.round(v), -127, 127).to(torch.int8)
    return out, truth
