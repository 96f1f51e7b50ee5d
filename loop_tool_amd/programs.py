"""The benchmark programs, written with the lazy frontend, plus their seeded
synthetic inputs.

Every builder takes a `capi.Library` (so the same program can be built with
any implementation of the C ABI) and returns (output tensor, input shapes);
the IR inputs are the leaf tensors in creation order.

Inputs follow a counter-based generator so every implementation sees the same
bits: SplitMix64 over (seed, element index); uniform [-1, 1) from the top 24
bits; normal by Box-Muller in float64, rounded to float32.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n):
    """n outputs of SplitMix64 seeded with `seed` (state += golden per draw)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, shape):
    n = int(np.prod(shape))
    x = splitmix64(seed, n)
    u = (x >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
    return (u * 2.0 - 1.0).astype(np.float32).reshape(shape)


def normal(seed, shape, std=1.0):
    n = int(np.prod(shape))
    m = (n + 1) // 2
    x = splitmix64(seed, 2 * m)
    u1 = ((x[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    u2 = (x[1::2] >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m, np.float64)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return (z[:n] * std).astype(np.float32).reshape(shape)


# ---------------------------------------------------------------- programs --
# Every builder takes `data`: None for size-only leaves (the program is then
# compiled and run explicitly on buffers), or one float32 array per input,
# bound to the leaves so that the lazy frontend can evaluate the program
# itself (`Tensor.numpy()`).
def _leaf(L, shape, data, i):
    if data is None:
        return L.Tensor(*shape)
    a = np.ascontiguousarray(data[i], np.float32).reshape(shape)
    return L.Tensor(data=a)


def matmul(L, M, N, K, data=None):
    """C[m,n] = sum_k A[m,k] * B[k,n]."""
    m, n, k = L.Symbol("M"), L.Symbol("N"), L.Symbol("K")
    A = _leaf(L, (M, K), data, 0)
    B = _leaf(L, (K, N), data, 1)
    C = (A.as_(m, k) * B.as_(k, n)).sum(k)
    return C, [(M, K), (K, N)]


def bmm(L, Bn, M, N, K, data=None):
    """C[b,m,n] = sum_k A[b,m,k] * B[b,k,n]: a batch of matmuls sharing b."""
    b, m, n, k = L.Symbol("B"), L.Symbol("M"), L.Symbol("N"), L.Symbol("K")
    A = _leaf(L, (Bn, M, K), data, 0)
    B = _leaf(L, (Bn, K, N), data, 1)
    C = (A.as_(b, m, k) * B.as_(b, k, n)).sum(k)
    return C, [(Bn, M, K), (Bn, K, N)]


def softmax(L, R, Cn, data=None):
    """Row softmax over C: max-reduce, subtract, exp, sum-reduce, divide."""
    r, c = L.Symbol("R"), L.Symbol("C")
    X = _leaf(L, (R, Cn), data, 0).as_(r, c)
    mx = X.max_reduce(c)
    e = (X - mx).exp()
    Y = e / e.sum(c)
    return Y, [(R, Cn)]


def conv2d(L, N, Ci, Co, H, W, k=3, pad=1, stride=1, data=None):
    """Y[n,co,ho,wo] = sum_{ci,kh,kw} Xpad[n,ci,ho*s+kh,wo*s+kw] * W[co,ci,kh,kw]
    with Xpad the zero-padded view of X (a pad nest plus a window sum)."""
    n, ci, h, w = L.Symbol("N"), L.Symbol("Ci"), L.Symbol("H"), L.Symbol("W")
    co, kh, kw = L.Symbol("Co"), L.Symbol("Kh"), L.Symbol("Kw")
    X = _leaf(L, (N, Ci, H, W), data, 0).as_(n, ci, h, w)
    Wt = _leaf(L, (Co, Ci, k, k), data, 1).as_(co, ci, kh, kw)
    if pad:
        X = X.pad(h, pad).pad(w, pad)
        sh, sw = X.shape[2], X.shape[3]
    else:
        sh, sw = h, w
    Y = L.convolve(X, Wt, [sh, sw], [kh, kw], stride)
    # convolve yields [N, Ho, Wo, Co]; lay the output out as NCHW
    s = Y.shape
    Y = Y.transpose(s[0], s[3], s[1], s[2])
    return Y, [(N, Ci, H, W), (Co, Ci, k, k)]


def maxpool2d(L, N, C, H, W, k=2, stride=2, data=None):
    n, c, h, w = L.Symbol("N"), L.Symbol("C"), L.Symbol("H"), L.Symbol("W")
    X = _leaf(L, (N, C, H, W), data, 0).as_(n, c, h, w)
    return L.maxpool(X, [h, w], k, stride), [(N, C, H, W)]


def pad1d(L, N, pre, post, data=None):
    n = L.Symbol("N")
    X = _leaf(L, (N,), data, 0).as_(n)
    return X.pad(n, pre, post), [(N,)]


def pad2d(L, N, C, H, W, pad=1, data=None):
    """Zero padding of the two spatial dims of an NCHW tensor (the pad nest
    of a convolution's input, materialised)."""
    n, c, h, w = L.Symbol("N"), L.Symbol("C"), L.Symbol("H"), L.Symbol("W")
    X = _leaf(L, (N, C, H, W), data, 0).as_(n, c, h, w)
    return X.pad(h, pad).pad(w, pad), [(N, C, H, W)]


def concat2d(L, B, M0, M1, data=None):
    b, m0, m1 = L.Symbol("N"), L.Symbol("M0"), L.Symbol("M1")
    A = _leaf(L, (B, M0), data, 0).as_(b, m0)
    Bt = _leaf(L, (B, M1), data, 1).as_(b, m1)
    return A | Bt, [(B, M0), (B, M1)]


def transpose2d(L, M, N, data=None):
    m, n = L.Symbol("M"), L.Symbol("N")
    A = _leaf(L, (M, N), data, 0).as_(m, n)
    return A.to([n, m]), [(M, N)]


def add_broadcast(L, M, N, data=None):
    m, n = L.Symbol("M"), L.Symbol("N")
    A = _leaf(L, (M, N), data, 0).as_(m, n)
    b = _leaf(L, (N,), data, 1).as_(n)
    return A + b, [(M, N), (N,)]


def conv_sample_indices(N, Co, H, W, n_random=65536, seed=12):
    """(n, co, h, w) output indices for a sampled float64 check of a
    convolution: every border pixel of the planes n in {0, N-1} for every
    output channel, plus n_random outputs drawn uniformly."""
    hh, ww = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    border = (hh == 0) | (hh == H - 1) | (ww == 0) | (ww == W - 1)
    bh, bw = hh[border], ww[border]
    idx = []
    for n in sorted({0, N - 1}):
        for co in range(Co):
            idx.append(np.stack([np.full_like(bh, n), np.full_like(bh, co), bh, bw], 1))
    rng = np.random.RandomState(seed)
    idx.append(np.stack([rng.randint(0, N, n_random), rng.randint(0, Co, n_random),
                         rng.randint(0, H, n_random), rng.randint(0, W, n_random)], 1))
    return np.concatenate(idx)


def conv_f64_at(X, Wt, idx, pad=1, chunk=4096):
    """float64 stride-1 convolution outputs Y[n, co, h, w] at the given
    indices (closed form: a dot product over channels and window taps)."""
    Xp = np.pad(X.astype(np.float64), ((0, 0), (0, 0), (pad, pad), (pad, pad)))
    Wd = Wt.astype(np.float64)
    k = Wt.shape[2]
    ci = np.arange(X.shape[1])
    out = np.empty(len(idx))
    for s in range(0, len(idx), chunk):
        n, co, h, w = idx[s:s + chunk].T
        acc = np.zeros(len(n))
        for a in range(k):
            for b in range(k):
                xs = Xp[n[:, None], ci[None, :], (h + a)[:, None], (w + b)[:, None]]
                acc += np.einsum("sc,sc->s", xs, Wd[co, :, a, b])
        out[s:s + chunk] = acc
    return out


def round_bf16(a):
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def reference_output(name, inputs, **kw):
    """float64 evaluation of a program (closed forms, not the oracle)."""
    x = [np.asarray(a, np.float64) for a in inputs]
    if name in ("matmul", "bmm"):
        return x[0] @ x[1]
    if name == "softmax":
        e = np.exp(x[0] - x[0].max(axis=1, keepdims=True))
        return e / e.sum(axis=1, keepdims=True)
    if name == "conv2d":
        X, W = x
        pad, stride = kw.get("pad", 1), kw.get("stride", 1)
        Xp = np.pad(X, ((0, 0), (0, 0), (pad, pad), (pad, pad)))
        k = W.shape[2]
        Ho = (Xp.shape[2] - k) // stride + 1
        Wo = (Xp.shape[3] - k) // stride + 1
        Y = np.zeros((X.shape[0], W.shape[0], Ho, Wo))
        for a in range(k):
            for b in range(k):
                patch = Xp[:, :, a:a + stride * Ho:stride, b:b + stride * Wo:stride]
                Y += np.einsum("nchw,oc->nohw", patch, W[:, :, a, b])
        return Y
    if name == "pad2d":
        p = kw.get("pad", 1)
        return np.pad(x[0], ((0, 0), (0, 0), (p, p), (p, p)))
    if name == "concat2d":
        return np.concatenate([x[0], x[1]], axis=1)
    if name == "transpose2d":
        return x[0].T
    if name == "maxpool2d":
        X = x[0]
        k, s = kw.get("k", 2), kw.get("stride", 2)
        Ho = (X.shape[2] - k) // s + 1
        Wo = (X.shape[3] - k) // s + 1
        out = np.full(X.shape[:2] + (Ho, Wo), -np.inf)
        for a in range(k):
            for b in range(k):
                out = np.maximum(out, X[:, :, a:a + s * Ho:s, b:b + s * Wo:s])
        return out
    raise KeyError(name)
