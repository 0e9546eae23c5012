"""Sub-integration bounds across blocks (dspx_timediv_*, restating
TimeDivide.C:36-506 and the loop of Subint::transformation, Subint.h:
253-330) against the oracle's restatement on sequences of blocks, and both
against what the division fixes: every sample of a contiguous stream lands
in exactly one division, divisions are numbered consecutively, each edge is
the nearest sample to start + k * interval, and a whole number of seconds
per division starts the divisions on a multiple of it within the day."""
import math

import numpy as np
import pytest

from oracle import timedivide as OT


def _stream(td_lib, td_or, start, rate, sizes, gaps=None):
    """Blocks of the given sizes from `start` (gaps[i] samples skipped
    before block i); returns both implementations' bounds per block."""
    day, sec, frac = start
    pos = 0
    out = []
    for i, n in enumerate(sizes):
        if gaps is not None:
            pos += gaps[i]
        t = frac + pos / rate
        bs = (day, sec + int(math.floor(t)), t - math.floor(t))
        a = td_lib.block(bs, rate, n)
        b = td_or.block(bs, rate, n)
        out.append((pos, n, a, b))
        pos += n
    return out


CASES = [
    # start (day, sec, frac), rate (Hz), seconds, turns, period
    ((60123, 4567, 0.123456789), 1.5625e6, 0.75, 0, 0),
    ((60123, 86399, 0.9), 3.125e6, 0.1, 0, 0),          # crosses midnight
    ((59000, 10, 0.25), 1e6, 1.0, 0, 0),                 # whole seconds: start rounded down
    ((59000, 7, 0.5), 64e6 / 3.0, 0.003, 0, 0),          # non-integer sample time
    ((60500, 300, 0.0), 1.6e7, 0, 4, 0.0893300),         # turns of a constant period
    ((60500, 300, 0.0), 2.5e5, 0, 1, 0.253),
]


@pytest.mark.parametrize("start,rate,seconds,turns,period", CASES)
def test_matches_oracle_contiguous_and_gapped(dspx, start, rate, seconds, turns, period):
    rng = np.random.default_rng(int(rate) % 1000 + int(seconds * 1000))
    interval = seconds if seconds else turns * period
    per_div = interval * rate
    sizes = [int(x) for x in rng.integers(max(1, int(per_div / 7)), int(per_div * 2.5) + 2, 30)]
    for gaps in (None, [int(g) for g in (rng.random(30) < 0.3) * rng.integers(1, int(per_div * 3) + 2, 30)]):
        a = dspx.TimeDivide(start, seconds, turns, period)
        b = OT.TimeDivide(start, seconds, turns, period)
        for pos, n, ga, gb in _stream(a, b, start, rate, sizes, gaps):
            assert ga == gb, (pos, n, ga, gb)
        a.close()


@pytest.mark.parametrize("start,rate,seconds,turns,period", CASES)
def test_contiguous_stream_partition(dspx, start, rate, seconds, turns, period):
    """Contiguous blocks: the divisions tile the stream without gaps or
    overlaps, are numbered consecutively, and each edge is the sample
    nearest start_time + k * interval (the start rounded down to a whole
    multiple of the division within the day when that is whole seconds)."""
    rng = np.random.default_rng(7)
    interval = seconds if seconds else turns * period
    per_div = interval * rate
    sizes = [int(x) for x in rng.integers(max(1, int(per_div / 5)), int(per_div * 3) + 2, 25)]
    td = dspx.TimeDivide(start, seconds, turns, period)
    day, sec, frac = start
    if seconds and seconds == int(seconds):
        t0 = (sec // int(seconds)) * int(seconds) - sec - frac  # start time relative to the stream start
    else:
        t0 = 0.0
    covered = []
    pos = 0
    for n in sizes:
        t = frac + pos / rate
        bounds = td.block((day, sec + int(math.floor(t)), t - math.floor(t)), rate, n)
        for b in bounds:
            covered.append((pos + b["idat_start"], pos + b["idat_start"] + b["ndat"], b["division"]))
        pos += n
    # no gaps, no overlaps, from the first sample to the last
    assert covered[0][0] == 0 and covered[-1][1] == pos
    for (s0, e0, d0), (s1, e1, d1) in zip(covered, covered[1:]):
        assert e0 == s1
        assert d1 in (d0, d0 + 1)
        if d1 == d0 + 1:
            # the edge of division d1: nearest sample to t0 + d1 * interval
            want = (t0 + d1 * interval) * rate
            assert abs(s1 - want) <= 0.5 + 1e-6, (s1, want)
    td.close()


def test_closed_form_whole_seconds(dspx):
    """1 MHz, 1 s divisions, stream starting 0.25 s into second 100 of the
    day: the start time is rounded down to second 100, so the first
    division holds 750000 samples and every later one 1000000."""
    td = dspx.TimeDivide((59000, 100, 0.25), 1.0)
    b = td.block((59000, 100, 0.25), 1e6, 3_500_000)
    assert [(x["idat_start"], x["ndat"], x["division"]) for x in b] == [
        (0, 750_000, 0), (750_000, 1_000_000, 1), (1_750_000, 1_000_000, 2), (2_750_000, 750_000, 3)]
    assert [x["end_reached"] for x in b] == [1, 1, 1, 0]
    assert all(x["division_ndat"] == 1_000_000 for x in b)
    # the next block continues division 3 to its end
    b = td.block((59000, 103, 0.75), 1e6, 600_000)
    assert [(x["idat_start"], x["ndat"], x["division"], x["end_reached"], x["new_division"]) for x in b] == [
        (0, 250_000, 3, 1, 0), (250_000, 350_000, 4, 0, 1)]
    td.close()


def test_gap_starts_a_new_division(dspx):
    """A block that starts after the current division's end starts the
    division holding its first sample (set_bounds :152-175)."""
    td = dspx.TimeDivide((59000, 0, 0.0), 2.0)
    td.block((59000, 0, 0.0), 1e3, 1500)
    b = td.block((59000, 9, 0.5), 1e3, 1000)
    assert [(x["idat_start"], x["ndat"], x["division"], x["new_division"]) for x in b] == [
        (0, 500, 4, 1), (500, 500, 5, 1)]
    td.close()


def test_reference_epoch_and_errors(dspx):
    td = dspx.TimeDivide((59000, 100, 0.25), 1.0, reference_epoch=(59000, 99, 0.5))
    b = td.block((59000, 100, 0.25), 1e6, 2_000_000)
    assert [(x["idat_start"], x["ndat"], x["division"]) for x in b] == [
        (0, 250_000, 0), (250_000, 1_000_000, 1), (1_250_000, 750_000, 2)]
    td.close()
    with pytest.raises(dspx.DspxError):
        dspx.TimeDivide((59000, 0, 0.0), 1.0, 2.0)
    with pytest.raises(dspx.DspxError):
        dspx.TimeDivide((59000, 0, 0.0), 0.0, 2.0)  # turns without a period
