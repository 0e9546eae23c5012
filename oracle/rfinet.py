"""CPU restatement of the RFI proposal network (test infrastructure only:
loaded by tests/ and bench.py's checks, never by the library).

features(): the per-(block, channel) features the HIP kernel k_rfi_*
computes from SK sums; bf16(): round-to-nearest-even to BF16; forward() and
loss_grad(): the 4-layer 1-D convolution across channels with ReLU and the
class-weighted logistic loss, in float64, with its gradient by backward
propagation (pinned by finite differences in tests/test_oracle_rfinet.py).
"""
import numpy as np

F, W, K, L = 4, 32, 19, 4
P = K // 2
SHAPES = [(W, F), (W, W), (W, W), (1, W)]  # (cout, cin) per layer


def nparam():
    return sum(co * ci * K + co for co, ci in SHAPES)


def unflatten(theta):
    out, o = [], 0
    for co, ci in SHAPES:
        w = theta[o:o + co * ci * K].reshape(co, ci, K)
        o += co * ci * K
        b = theta[o:o + co]
        o += co
        out.append((w, b))
    return out


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def features(sums, counts, Nd=1):
    """sums [nblk][nchan][npol][2] float32, counts [nblk][nchan] ->
    feat [nblk][4][nchan] float32 (float32 arithmetic as the kernel's)."""
    sums = np.asarray(sums, np.float32)
    counts = np.asarray(counts, np.uint32)
    nblk, nchan, npol, _ = sums.shape
    lp = np.zeros((nblk, nchan), np.float32)
    v1 = np.zeros((nblk, nchan), np.float32)
    vm = np.zeros((nblk, nchan), np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        for b in range(nblk):
            for c in range(nchan):
                Mp = int(counts[b, c])
                if Mp <= 2:
                    continue
                mfac = np.float32(np.float32(Mp * Nd + 1) / np.float32(Mp - 1))
                s1 = np.float32(0)
                vs = np.float32(0)
                vmax = np.float32(0)
                for p in range(npol):
                    S1, S2 = sums[b, c, p]
                    if S1 > 0:
                        V = np.float32(mfac * np.float32(np.float32(Mp) * np.float32(S2 / np.float32(S1 * S1)) - 1))
                    else:
                        V = np.float32(1)
                    s1 = np.float32(s1 + S1)
                    vs = np.float32(vs + np.float32(V - 1))
                    vmax = max(vmax, abs(np.float32(V - 1)))
                lp[b, c] = np.log(np.float32(s1 / np.float32(Mp)) + np.float32(1e-20))
                v1[b, c] = vs
                vm[b, c] = vmax
    v1 = (v1 / np.float32(npol)).astype(np.float32)
    mean = lp.astype(np.float64).mean(axis=0).astype(np.float32)
    feat = np.zeros((nblk, F, nchan), np.float32)
    feat[:, 0] = lp - mean
    feat[:, 1] = v1
    feat[:, 2] = vm
    feat[1:, 3] = lp[1:] - lp[:-1]
    return feat


def _conv(a, w, b):
    """a [nblk][cin][nchan], w [cout][cin][K], b [cout] -> [nblk][cout][nchan]
    (zero padding, 'same')."""
    nblk, cin, n = a.shape
    ap = np.zeros((nblk, cin, n + K - 1), a.dtype)
    ap[:, :, P:P + n] = a
    cols = np.stack([ap[:, :, k:k + n] for k in range(K)], axis=2)  # [nblk][cin][K][n]
    return np.einsum("oik,bikn->bon", w, cols) + b[None, :, None]


def forward(theta, feat):
    """Logits [nblk][nchan] and the activations (float64)."""
    layers = unflatten(np.asarray(theta, np.float64))
    acts = [np.asarray(feat, np.float64)]
    for l, (w, b) in enumerate(layers):
        z = _conv(acts[-1], w, b)
        acts.append(np.maximum(z, 0) if l < L - 1 else z)
    return acts[-1][:, 0], acts


def loss_grad(theta, feat, labels, pos_weight=1.0):
    """Mean class-weighted logistic loss and its gradient (float64)."""
    layers = unflatten(np.asarray(theta, np.float64))
    z, acts = forward(theta, feat)
    y = np.asarray(labels, np.float64)
    wy = np.where(y > 0, pos_weight, 1.0)
    n = z.size
    sp = np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z)))
    loss = float(np.sum(wy * (sp - y * z)) / n)
    g = (wy * (1 / (1 + np.exp(-z)) - y) / n)[:, None, :]
    grads = [None] * L
    for l in range(L - 1, -1, -1):
        w, b = layers[l]
        a = acts[l]
        nblk, cin, nn = a.shape
        ap = np.zeros((nblk, cin, nn + K - 1))
        ap[:, :, P:P + nn] = a
        cols = np.stack([ap[:, :, k:k + nn] for k in range(K)], axis=2)
        gw = np.einsum("bon,bikn->oik", g, cols)
        gb = g.sum(axis=(0, 2))
        grads[l] = (gw, gb)
        if l > 0:
            gp = np.zeros((nblk, w.shape[0], nn + K - 1))
            gp[:, :, P:P + nn] = g
            # dA[i][x] = sum_o sum_k w[o][i][k] g[o][x - k + P]
            da = np.zeros((nblk, cin, nn))
            for k in range(K):
                da += np.einsum("oi,bon->bin", w[:, :, k], gp[:, :, K - 1 - k:K - 1 - k + nn])
            g = da * (acts[l] > 0)
    flat = []
    for gw, gb in grads:
        flat.append(gw.ravel())
        flat.append(gb.ravel())
    return loss, np.concatenate(flat)
