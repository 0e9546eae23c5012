"""The oracle's SpectralKurtosis with several resolutions (adaptive SK,
SpectralKurtosis.C:849-972, 1105-1462) pinned on the CPU:

- one resolution reduces to the single-resolution restatement (every
  detector, with and without omit_outliers and overlap);
- the time-scrunched sums equal sums formed directly from the samples of
  the blocks the reference adds (blocks (b + i) * offset of the finer
  resolution): where that offset equals the finer noverlap this is the
  coarser resolution's own blocks, checked against the oracle's compute at
  the coarser M; elsewhere (overlapping finer blocks) the reference's stride
  is kept, checked against numpy sums over those samples;
- a coarser block's zap flags the finest resolution's rows it spans: a
  channel modulated at a period no 64-sample block sees is caught only by
  the 1024-sample resolution, and its flagged rows are those rows;
- detect_tscr takes the last resolution's blocks and flags the mask's first
  rows, as many as that resolution has blocks;
- an incompatible pair (the reference's tscrunch would read past its input)
  is an error;
- the mask starts at the coarsest resolution's first output sample and
  zeroes the finest blocks of flagged rows."""
import os

import numpy as np
import pytest

from helpers import sk_rfi_input

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sk_limits.npz")


@pytest.fixture(scope="module")
def rtab():
    z = np.load(GOLDEN)
    return {M: (z["rlower%d" % M], z["rupper%d" % M], z["rsym%d" % M]) for M in (64, 128, 256, 512, 1024)}


def _res(rtab, spec):
    return [(M, nov) + rtab[M] + (None,) for M, nov in spec]


@pytest.mark.parametrize("flags", [1, 2, 3])
@pytest.mark.parametrize("nov,omit", [(1, False), (1, True), (4, True)])
def test_one_resolution_is_the_single_restatement(oracle, rtab, flags, nov, omit):
    M, nchan = 64, 12
    x = sk_rfi_input(nchan, 2, 40, M, seed=3, tone_chan=4, impulse_blk=7)
    ndat = x.shape[2]
    sums, counts = oracle.sk_compute(x, M, noverlap=nov)
    lo, up, sy = rtab[M]
    m1, s1, c1 = oracle.sk_detect(sums, counts, M, lo, up, sy, flags=flags, omit_outliers=omit, noverlap=nov)
    m2, s2, c2 = oracle.sk_detect_res(sums, counts, ndat, _res(rtab, [(M, nov)]), flags=flags, omit_outliers=omit)
    assert np.array_equal(m1, m2) and np.array_equal(s1, s2) and np.array_equal(c1, c2)
    assert m1.any()


def test_tscrunch_equals_coarser_blocks(oracle, rtab):
    """(64, noverlap 1) -> (256, noverlap 4): offset 1 = the finer noverlap,
    so coarse block b sums the four consecutive 64-sample blocks of samples
    [64 b, 64 b + 256) -- the 256-sample blocks every 64 samples themselves."""
    nchan = 5
    x = sk_rfi_input(nchan, 2, 48, 64, seed=8)
    ndat = x.shape[2]
    sums, counts = oracle.sk_compute(x, 64)
    _, ts, tc = oracle.sk_detect_res(sums, counts, ndat, _res(rtab, [(64, 1), (256, 4)]), flags=0)
    direct, dc = oracle.sk_compute(x, 256, noverlap=4)
    n1 = direct.shape[0]
    assert np.array_equal(tc[:n1], dc)
    np.testing.assert_allclose(ts[:n1], direct, rtol=2e-6)


def test_tscrunch_keeps_the_reference_stride(oracle, rtab):
    """(128, noverlap 2) -> (512, noverlap 8): offset (512/8)/(128/2) = 1
    while the finer blocks overlap (noverlap 2), so coarse block b adds the
    overlapping finer blocks b..b+3 (samples [64(b+i), 64(b+i) + 128)), as
    tscrunch_sums's stride does; numpy sums over exactly those samples."""
    nchan, M0 = 4, 128
    x = sk_rfi_input(nchan, 2, 30, M0, seed=9)
    ndat = x.shape[2]
    sums, counts = oracle.sk_compute(x, M0, noverlap=2)
    _, ts, tc = oracle.sk_detect_res(sums, counts, ndat, _res(rtab, [(128, 2), (512, 8)]), flags=0)
    n1, _ = oracle.sk_blocks(ndat, 512, 8)
    sq = (np.abs(x.astype(np.complex128)) ** 2)
    for b in range(0, n1, 7):
        for c in range(nchan):
            for p in range(2):
                v = np.concatenate([sq[c, p, 64 * (b + i):64 * (b + i) + 128] for i in range(4)])
                np.testing.assert_allclose(ts[b, c, p], [v.sum(), (v * v).sum()], rtol=1e-5)
            assert tc[b, c] == 4 * 128


def test_coarse_zap_flags_the_rows_it_spans(oracle, rtab):
    """A channel whose power switches between two levels every 256 samples
    (Gaussian within each 64-sample block) is flagged by the coarse
    detect_skfb, not the 64-sample one; the rows a zapped block of the last
    resolution (1024 samples, noverlap 16: block b from sample 64 b) flags
    are the 64-sample rows b .. b + 15 it spans."""
    nchan, nblk = 6, 64
    rng = np.random.default_rng(4)
    n = nblk * 64
    x = (rng.standard_normal((nchan, 2, n)) + 1j * rng.standard_normal((nchan, 2, n))) * 10
    mod = np.where((np.arange(n) // 256) % 2 == 0, 1.0, 3.0)
    x[2] *= mod
    x = x.astype(np.complex64)
    sums, counts = oracle.sk_compute(x, 64)
    fine, _, _ = oracle.sk_detect_res(sums, counts, n, _res(rtab, [(64, 1)]), flags=1)
    mask, ts, tc = oracle.sk_detect_res(sums, counts, n, _res(rtab, [(64, 1), (256, 4), (1024, 16)]), flags=1)
    assert fine[:, 2].mean() < 0.1
    # the coarse estimator over the last resolution's blocks, from the
    # time-scrunched sums, and the rows each zapped block spans
    n2, _ = oracle.sk_blocks(n, 1024, 16)
    lo, up, _ = rtab[1024]
    expect = fine.copy()
    zapped = []
    for b in range(n2):
        Mp = int(tc[b, 2])
        S1, S2 = ts[b, 2, :, 0], ts[b, 2, :, 1]
        V = np.float32((Mp + 1) / (Mp - 1)) * (Mp * (S2 / (S1 * S1)) - 1)
        if np.any(V > up[Mp]) or np.any(V < lo[Mp]):
            zapped.append(b)
            # flag_offset = nflag * flag_step / noverlap = 16 * 1 / 16: rows b + k, k < 16
            expect[b:b + 16, 2] = 1
    assert len(zapped) > n2 // 2
    got = mask[:, 2].astype(bool)
    assert np.all(got[expect[:, 2].astype(bool)])
    assert got.mean() > 0.5


def test_tscr_flags_the_first_rows_of_the_last_resolution(oracle, rtab):
    """detect_tscr zaps a tone channel; with resolutions (64, 1) -> (256, 4)
    it flags that channel's first npart(256, 4) rows of the 64-sample mask
    (the reference's loop over the last resolution's blocks)."""
    nchan, nblk = 8, 40
    x = sk_rfi_input(nchan, 2, nblk, 64, seed=12, tone_chan=3)
    n = x.shape[2]
    sums, counts = oracle.sk_compute(x, 64)
    mask, _, _ = oracle.sk_detect_res(sums, counts, n, _res(rtab, [(64, 1), (256, 4)]), flags=4,
                                      tscr_limits={t: (0.8, 1.2) for t in range(257, 32768)})
    n1, _ = oracle.sk_blocks(n, 256, 4)
    assert n1 < nblk
    assert mask[:n1, 3].all() and not mask[n1:, 3].any()
    assert not mask[:, [0, 1, 2, 4, 5, 6, 7]].any()


def test_incompatible_resolutions_are_an_error(oracle, rtab):
    """(64, 1) -> (256, 1): offset 4 = nsum, so the last coarse block would
    read finer block (npart - 1) * 4 + 12, past the input (the reference
    asserts)."""
    x = sk_rfi_input(3, 2, 32, 64, seed=1)
    sums, counts = oracle.sk_compute(x, 64)
    with pytest.raises(ValueError):
        oracle.sk_detect_res(sums, counts, x.shape[2], _res(rtab, [(64, 1), (256, 1)]), flags=1)


def test_mask_with_resolutions(oracle, rtab):
    """The output starts at the coarsest resolution's first sample (M - M /
    noverlap = 192 for (256, 4)), in finest blocks: rows nskip = 3 on, each
    64-sample block of the output zeroed where its row is flagged."""
    nchan, nblk, M = 4, 30, 64
    x = sk_rfi_input(nchan, 2, nblk, M, seed=2)
    n = x.shape[2]
    rng = np.random.default_rng(0)
    mask = (rng.random((nblk, nchan)) < 0.3).astype(np.uint8)
    out, nout = oracle.sk_mask_res(x, 64, 1, 256, 4, mask)
    n1, nout1 = oracle.sk_blocks(n, 256, 4)
    assert nout == nout1 == n1 * 64 - 192
    for c in range(nchan):
        for b in range(nout // M):
            seg = out[c, :, b * M:(b + 1) * M]
            if mask[b + 3, c]:
                assert not seg.any()
            else:
                assert np.array_equal(seg, x[c, :, 192 + b * M:192 + (b + 1) * M])
