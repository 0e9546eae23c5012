"""Inter-channel dispersion delays (dspsr -K): SampleDelay with the
dispersion delay function, the fractional delay carried by the chirp, and
their combined effect on a dispersed broadband impulse (CPU: oracle and the
library's host-side delays)."""
import math

import numpy as np
import pytest


def test_sample_delays_invariants(oracle):
    """The integer delays are the rounded dispersion delays relative to the
    band centre (zero at the centre channel, the remainder within half a
    sample), the delay difference between two channels is the difference of
    the cold-plasma delays, and delay + applied offset is the same for every
    channel (which is what aligns them)."""
    cfreq, bw, nchan, dm = 1400.0, 64.0, 31, 120.0
    sd = oracle.sample_delays(cfreq, bw, nchan, dm)
    d, f, a = sd["delay"], sd["frac"], sd["applied"]
    assert d[nchan // 2] == 0 and f[nchan // 2] == 0.0
    assert np.all(np.abs(f) <= 0.5)
    lib = oracle.lib()
    rate = abs(bw) / nchan * 1e6
    fc = [lib.ox_channel_centre(cfreq, bw, nchan, 0, c) for c in range(nchan)]
    # ox_delay_time is the geometry's cold-plasma delay (seconds) at a frequency
    for c in (0, 7, 30):
        dt = (lib.ox_delay_time(dm, fc[nchan // 2]) - lib.ox_delay_time(dm, fc[c])) * rate
        assert abs((d[c] + f[c]) - dt) < 1e-6 * max(1.0, abs(dt))
    # higher frequencies arrive earlier: the delay relative to the centre grows with frequency
    assert np.all(np.diff(d + f) > 0)
    assert len(set((a + d).tolist())) == 1
    assert a.min() >= 0 and sd["total"] == a.max()
    assert sd["zero_delay"] == (a + d)[0]  # the span-centre delay is 0 for this band


def test_library_sample_delays_match_oracle(dspx, oracle):
    for cfreq, bw, nchan, dm in [(1400.0, 64.0, 31, 120.0), (800.0, 200.0, 256, 35.5), (1405.0, 16.0, 1, 56.77),
                                 (430.0, -20.0, 64, 12.0)]:
        a = dspx.sample_delays(cfreq, bw, nchan, dm)
        b = oracle.sample_delays(cfreq, bw, nchan, dm)
        for k in ("delay", "applied"):
            assert np.array_equal(a[k], b[k])
        assert np.array_equal(a["frac"], b["frac"])
        assert a["zero_delay"] == b["zero_delay"] and a["total"] == b["total"]


def dispersed_impulse(oracle, cfreq, bw, nchan, dm, ndat, t0):
    """A broadband impulse at sample t0 of the band centre, seen in each
    channel: the cold-plasma delay of the channel centre relative to the band
    centre (-(delay + frac) samples, fractional part included) and the
    in-channel dispersion (the conjugate of the FP64 chirp without delay
    term), synthesised on an ndat-point grid."""
    sd = oracle.sample_delays(cfreq, bw, nchan, dm)
    H = oracle.response(cfreq, bw, nchan, dm, ndat, "f64") * math.sqrt(ndat)
    k = np.fft.fftfreq(ndat) * ndat  # signed bin index, FFT order
    x = np.zeros((nchan, 1, ndat), np.complex128)
    for c in range(nchan):
        s = sd["delay"][c] + sd["frac"][c]
        spec = np.exp(-2j * np.pi * k * t0 / ndat) * np.conj(H[c]) * np.exp(2j * np.pi * k * s / ndat)
        x[c, 0] = np.fft.ifft(spec)
    return x, sd


@pytest.mark.parametrize("interchan", [True, False])
def test_interchan_dedispersion_aligns_impulse(oracle, interchan):
    """After convolution with the fractional-delay chirp and the integer
    offsets, the impulse is one sample at the same output index in every
    channel.  Without the fractional term the channels whose delay is far
    from an integer smear it."""
    cfreq, bw, nchan, dm = 1400.0, 16.0, 8, 10.0
    g = oracle.geometry(cfreq, bw, nchan, dm)
    n = g["nfft"]
    ndat = 4 * n
    t0 = ndat // 2
    x, sd = dispersed_impulse(oracle, cfreq, bw, nchan, dm, ndat, t0)
    resp = oracle.response(cfreq, bw, nchan, dm, n, 0, interchan=interchan)
    conv = oracle.convolve(x.astype(np.complex64), resp, g["impulse_pos"], g["impulse_neg"])
    nout = conv.shape[2] - sd["total"]
    out = oracle.retard(conv, sd["applied"], nout)
    y = np.abs(out[:, 0]) / math.sqrt(n)
    expect = t0 - g["impulse_pos"] - (sd["applied"][0] + sd["delay"][0])
    peaks = y.max(axis=1)
    if interchan:
        assert np.all(np.argmax(y, axis=1) == expect)
        assert peaks.min() > 0.99
        for c in range(nchan):
            assert np.delete(y[c], expect).max() < 0.02
    else:
        far = np.abs(sd["frac"]) > 0.3
        assert far.any()
        assert peaks[far].max() < 0.9
