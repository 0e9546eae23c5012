"""CPU restatement of dsp::TimeDivide's sub-integration bounds (test
infrastructure only: loaded by tests/, never by the library).

Follows Signal/Pulsar/TimeDivide.C: set_start_time (:36-67), set_bounds
(:124-338), set_boundaries with a division in seconds or in turns of a
constant period (:346-433) and the rounding of the division edges to input
samples (:467-506); and the loop of Subint::transformation
(Signal/Pulsar/dsp/Subint.h:253-330) that calls set_bounds until the block
is used up.

Epochs are (day, second, fraction) triples held as Python ints and a float,
settled after every sum and difference as psrchive's MJD does (a dependency
absent from the reference tree; its representation is restated here).
Rounding: rint / lrint round half to even, as Python's round() does.
"""
import math


class Epoch:
    __slots__ = ("day", "sec", "frac")

    def __init__(self, day=0, sec=0, frac=0.0):
        fs = math.floor(frac)
        sec += int(fs)
        frac -= fs
        ds, sec = divmod(sec, 86400)  # floor division: 0 <= sec < 86400
        self.day, self.sec, self.frac = day + ds, sec, frac

    def seconds(self):
        return float(self.day) * 86400.0 + float(self.sec) + self.frac

    def __add__(self, s):
        return Epoch(self.day, self.sec, self.frac + s)

    def __sub__(self, o):
        return Epoch(self.day - o.day, self.sec - o.sec, self.frac - o.frac)

    def __lt__(self, o):
        return (self - o).seconds() < 0.0

    def __gt__(self, o):
        return (self - o).seconds() > 0.0

    def is_zero(self):
        return self.day == 0 and self.sec == 0 and self.frac == 0.0


def _max(a, b):
    return b if a < b else a


def _min(a, b):
    return b if b < a else a


class TimeDivide:
    def __init__(self, start, seconds=0.0, turns=0.0, period=0.0, reference_epoch=None):
        assert (seconds > 0) != (turns > 0) and (turns == 0 or period > 0)
        self.seconds, self.turns, self.period = seconds, turns, period
        st = Epoch(*start)
        if reference_epoch is not None and not Epoch(*reference_epoch).is_zero():
            st = Epoch(*reference_epoch)
        elif seconds > 0 and seconds == int(seconds):
            s = int(seconds)
            st = Epoch(st.day, (st.sec // s) * s, 0.0)
        self.start = st
        self.lower = Epoch()
        self.upper = Epoch()
        self.current_end = Epoch()
        self.valid = False
        self.division = 0
        self.division_ndat = 0

    def _edges(self, m1, m2, in_start, rate):
        samples = round((m1 - in_start).seconds() * rate)
        self.lower = in_start + samples / rate
        secs = (m2 - self.lower).seconds()
        assert secs > 0
        self.division_ndat = round(secs * rate)
        self.upper = self.lower + self.division_ndat / rate

    def _division(self, t, in_start, rate):
        ds = _max(self.start, t)
        spi = self.seconds if self.seconds > 0 else self.period * self.turns
        self.division = int((ds - self.start).seconds() / spi)
        self._edges(self.start + self.division * spi, self.start + (self.division + 1) * spi, in_start, rate)

    def set_bounds(self, in_start, rate, ndat):
        """One set_bounds: dict(idat_start, ndat, division, division_ndat,
        new_division, end_reached, in_next, is_valid)."""
        s = Epoch(*in_start) if not isinstance(in_start, Epoch) else in_start
        end = s + ndat / rate
        ds = _max(self.current_end, s) if self.valid else s
        b = dict(idat_start=0, ndat=0, division=0, division_ndat=0, new_division=0, end_reached=0, in_next=0,
                 is_valid=0)
        if end < self.lower or ds + 0.5 / rate > self.upper:
            b["new_division"] = 1
            self._division(ds + 0.55 / rate, s, rate)
        ds = _max(self.lower, ds)
        i0 = round((ds - s).seconds() * rate)
        assert i0 >= 0
        b["division"], b["division_ndat"] = self.division, self.division_ndat
        if i0 >= ndat:
            self.valid = False
            return b
        de = _min(end, self.upper)
        i1 = round((de - s).seconds() * rate)
        if i1 <= i0:
            raise ValueError("idat_end <= idat_start")
        if i1 > ndat:
            i1 = ndat
        elif i1 < ndat:
            b["in_next"] = 1
        b["idat_start"], b["ndat"] = i0, i1 - i0
        if (self.upper - de).seconds() * rate < 0.5:
            b["end_reached"] = 1
        b["is_valid"] = 1
        self.valid = True
        self.current_end = s + i1 / rate
        return b

    def block(self, in_start, rate, ndat):
        """The set_bounds calls of Subint::transformation for one block: the
        valid bounds in order."""
        out = []
        while True:
            b = self.set_bounds(in_start, rate, ndat)
            if b["is_valid"]:
                out.append(b)
            if not b["in_next"]:
                return out
