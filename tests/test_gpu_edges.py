"""Edge cases of the convolution and the fused path on the GPU: every FFT
length the four-step split supports (2^12 .. 2^22, both forms) against the
oracle, a block too short for one FFT, and a fold histogram too large for
LDS (a clean error, nothing launched)."""
import math

import numpy as np
import pytest

from helpers import from_dev, rel_rms, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("m", list(range(12, 23)))
def test_convolution_every_length(ctx, dspx, oracle, m, form):
    import torch
    n = 1 << m
    pos, neg = 100 + m, 80 + 2 * m
    step = n - pos - neg
    ndat = n + step + step // 3  # two parts and a ragged tail
    rng = np.random.default_rng(m)
    resp = (np.exp(1j * rng.uniform(0, 2 * np.pi, (1, n))) / math.sqrt(n)).astype(np.complex64)
    x = ((rng.standard_normal((1, 2, ndat)) + 1j * rng.standard_normal((1, 2, ndat))) * 10).astype(np.complex64)
    expect = oracle.convolve(x, resp, pos, neg)
    nout = expect.shape[2]
    npart = nout // step
    assert npart == 2
    conv = dspx.Convolution(ctx, 1, 2, n, pos, neg, chirp_mode=dspx.CHIRP_TABLE, form=form)
    conv.set_response(resp)
    out = torch.zeros((1, 2, nout, 2), dtype=torch.float32, device="cuda")
    print("convolution length 2^%d form %d" % (m, form), flush=True)
    conv.perform(to_dev(x.view(np.float32)), 2 * ndat * 2, ndat * 2, ndat, out, 2 * nout * 2, nout * 2, npart)
    got = from_dev(out).view(np.complex64).reshape(1, 2, nout)
    e = rel_rms(got, expect)
    print("rel rms %.3g" % e)
    assert e < 1e-5
    conv.close()


def test_block_shorter_than_one_fft(ctx, dspx):
    """ndat < nfft: no parts, no output; execute and the filterbank return at
    once and leave the profiles and hits as they were."""
    import torch
    cfreq, bw, nchan, dm = 1400.0, 50.0, 16, 20.0
    g = dspx.geometry(cfreq, bw, nchan, dm)
    ndat = g["nfft"] - 1
    L = dspx.LoadToFold(ctx, 8, nchan, 2, ndat, cfreq, bw, dm, 256, state=dspx.INTENSITY, geom=g)
    assert L.npart == 0 and L.ndat_out == 0
    L.set_fold(1e-4, [(0, 0, 0.0, 0)])
    raw = torch.zeros(ndat * nchan * 2 * 2, dtype=torch.int8, device="cuda")
    prof = torch.full((1, nchan, 256), 7.0, device="cuda")
    hits = torch.full((1, 256), 3, dtype=torch.int32, device="cuda")
    print("empty block", flush=True)
    L.execute(raw, prof, hits)
    assert float(prof.min()) == 7.0 and float(prof.max()) == 7.0
    assert int(hits.min()) == 3 and int(hits.max()) == 3
    L.close()


def test_fold_histogram_too_large_is_an_error(ctx, dspx):
    """A fused fold whose LDS histogram (slots x values x bins) exceeds the
    workgroup's LDS is refused by set_fold with DSPX_E_UNSUPPORTED; the
    context stays usable."""
    cfreq, bw, nchan, dm = 1400.0, 50.0, 16, 20.0
    g = dspx.geometry(cfreq, bw, nchan, dm)
    ndat = 4 * g["nfft"]
    L = dspx.LoadToFold(ctx, 8, nchan, 2, ndat, cfreq, bw, dm, 16384, state=dspx.STOKES, ndim=4, geom=g)
    rate = bw / nchan * 1e6
    chunks = dspx.fold_chunks(L.ndat_out, rate, 0.01, g["impulse_pos"], nsub=2)
    print("histogram too large", flush=True)
    with pytest.raises(dspx.DspxError) as e:
        L.set_fold((1.0 / rate) / 0.01, chunks)
    assert e.value.code == dspx.E_UNSUPPORTED
    L.close()
    ctx.synchronize()
