"""CPU oracle of the learned prefilter (test infrastructure only).

MLP 16 -> 64 -> 64 -> 1 with ReLU and a sigmoid output, bf16 weights and
activations, fp32 accumulation in a fixed order (k = 0, 1, ... for every
dot product, samples s = 0, 1, ... for every batch sum), binary cross-entropy
with a positive-class weight, Adam.  The HIP kernels in
pastis_amd/csrc/k_filter.hip compute the same expressions in the same order.
There is no counterpart in the reference: the filter sits beside the k-mer
prefilter of src/DistPWRunner.cpp:48-54.

Parameter layout (flat fp32, 5313 values):
  W1[16][64] (in-major), b1[64], W2[64][64], b2[64], w3[64], b3[1]
"""
import numpy as np

NF, H = 16, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NF * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1
assert NPARAM == 5313


def bf16(x):
    """fp32 -> bf16 -> fp32, round to nearest even (finite inputs)."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(p):
    p = np.asarray(p, np.float32)
    return (p[OFF_W1:OFF_B1].reshape(NF, H), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3:NPARAM])


def init_params(seed=0):
    """He-uniform weights, zero biases (same draw as pa_filter_init)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.uniform(-1, 1, NF * H) * np.sqrt(6.0 / NF)
    p[OFF_W2:OFF_B2] = rng.uniform(-1, 1, H * H) * np.sqrt(6.0 / H)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, H) * np.sqrt(6.0 / H)
    return p


def _dot_seq(a, W, b):
    """a[B,K] @ W[K,N] + b, fp32 accumulation over k in order k = 0..K-1."""
    acc = np.zeros((a.shape[0], W.shape[1]), np.float32)
    for k in range(W.shape[0]):
        acc = (acc + a[:, k:k + 1] * W[k:k + 1, :]).astype(np.float32)
    return (acc + b).astype(np.float32)


def forward(params, x):
    """Returns (logit z3, dict of intermediates) for standardized features."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    xb = bf16(x)
    z1 = _dot_seq(xb, bf16(W1), b1)
    h1b = bf16(np.maximum(z1, 0))
    z2 = _dot_seq(h1b, bf16(W2), b2)
    h2b = bf16(np.maximum(z2, 0))
    z3 = _dot_seq(h2b, bf16(w3).reshape(H, 1), b3)[:, 0]
    return z3, dict(xb=xb, z1=z1, h1b=h1b, z2=z2, h2b=h2b)


def sigmoid(z):
    return (1.0 / (1.0 + np.exp(-z.astype(np.float64)))).astype(np.float32)


def loss_and_grad(params, x, y, pos_weight=1.0):
    """Mean weighted BCE and its gradient (flat, layout above)."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    z3, c = forward(params, x)
    p = sigmoid(z3)
    y = np.asarray(y, np.float32)
    B = len(y)
    eps = 1e-7
    loss = float(np.mean(-(pos_weight * y * np.log(np.maximum(p, eps)) +
                           (1 - y) * np.log(np.maximum(1 - p, eps)))))
    dz3 = (((1 - y) * p - pos_weight * y * (1 - p)) / B).astype(np.float32)
    dh2 = (dz3[:, None] * bf16(w3)[None, :]).astype(np.float32)
    dz2 = np.where(c["z2"] > 0, dh2, 0).astype(np.float32)
    W2b = bf16(W2)
    dh1 = np.zeros((B, H), np.float32)
    for k in range(H):   # dh1[:, i] = sum_k dz2[:, k] * W2[i, k], k in order
        dh1 = (dh1 + dz2[:, k:k + 1] * W2b[:, k][None, :]).astype(np.float32)
    dz1 = np.where(c["z1"] > 0, dh1, 0).astype(np.float32)

    def bsum(a, b):      # sum_s a[s, :, None] * b[s, None, :], s in order
        acc = np.zeros((a.shape[1], b.shape[1]), np.float32)
        for s in range(a.shape[0]):
            acc = (acc + a[s][:, None] * b[s][None, :]).astype(np.float32)
        return acc

    def vsum(a):
        acc = np.zeros(a.shape[1:], np.float32)
        for s in range(a.shape[0]):
            acc = (acc + a[s]).astype(np.float32)
        return acc

    g = np.zeros(NPARAM, np.float32)
    g[OFF_W1:OFF_B1] = bsum(c["xb"], dz1).ravel()
    g[OFF_B1:OFF_W2] = vsum(dz1)
    g[OFF_W2:OFF_B2] = bsum(c["h1b"], dz2).ravel()
    g[OFF_B2:OFF_W3] = vsum(dz2)
    g[OFF_W3:OFF_B3] = bsum(c["h2b"], dz3[:, None])[:, 0]
    g[OFF_B3] = vsum(dz3[:, None])[0]
    return loss, g


def adam_step(params, m, v, g, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    m[:] = (beta1 * m + (1 - beta1) * g).astype(np.float32)
    v[:] = (beta2 * v + (1 - beta2) * g * g).astype(np.float32)
    mh = m / np.float32(1 - beta1 ** t)
    vh = v / np.float32(1 - beta2 ** t)
    params -= (lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return params


def predict(params, x):
    return sigmoid(forward(params, x)[0])


# ---- features ---------------------------------------------------------------
# residue code (ARNDCQEGHILKMFPSTWYVBZX*J) -> murphy10 class in the numbering
# Alphabet::init() gives Murphy10ProteinAlph (inc/Alphabet.hpp:89-103,
# src/Alphabet.cpp:31-51): first appearance of each class representative.
MURPHY10 = np.array([0, 1, 2, 2, 3, 2, 2, 4, 5, 6, 6, 1, 6, 7, 8, 9, 9, 7, 7,
                     6, 10, 11, 12, 13, 14], np.uint32)
SKETCH, KMER = 16, 6
EMPTY_HASH = 0xFFFFFFFF


def mix32(x):
    """xorshift-multiply 32-bit mixer (uint32 arrays, wrapping)."""
    x = np.asarray(x, np.uint32).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint32(16)
        x *= np.uint32(0x7FEB352D)
        x ^= x >> np.uint32(15)
        x *= np.uint32(0x846CA68B)
        x ^= x >> np.uint32(16)
    return x


def kmer_codes(m):
    """Base-15 codes of the murphy10 6-mers of class array m."""
    m = np.asarray(m, np.uint64)
    if len(m) < KMER:
        return np.zeros(0, np.uint32)
    code = np.zeros(len(m) - KMER + 1, np.uint64)
    for t in range(KMER):
        code = code * 15 + m[t:len(m) - KMER + 1 + t]
    return code.astype(np.uint32)


def salts():
    with np.errstate(over="ignore"):
        return (np.uint32(0x9E3779B9) *
                np.arange(1, SKETCH + 1, dtype=np.uint32)).astype(np.uint32)


def seq_sketch(codes):
    """(composition fractions [10] fp32, MinHash [16] uint32, length)."""
    codes = np.asarray(codes, np.uint8)
    m = MURPHY10[codes] if len(codes) else np.zeros(0, np.uint32)
    std = m[m < 10]
    cnt = np.bincount(std, minlength=10)[:10].astype(np.float32)
    tot = np.float32(max(len(std), 1))
    comp = (cnt / tot).astype(np.float32)
    mh = np.full(SKETCH, EMPTY_HASH, np.uint32)
    km = kmer_codes(m)
    if len(km):
        for h, s in enumerate(salts()):
            mh[h] = mix32(km ^ s).min()
    return comp, mh, len(codes)


def _lg2(x):
    return np.float32(np.log2(np.float64(x)))


def pair_features(p, R, C):
    """16 fp32 features of pair p (fields r, c, count, kscore, seed[4]) from
    sketches R, C, with the operation order of the kernels."""
    f32 = np.float32
    (rcomp, rmh, lr), (ccomp, cmh, lc) = R, C
    Lr, Lc = f32(lr), f32(lc)
    Lr1, Lc1 = max(Lr, f32(1)), max(Lc, f32(1))
    mx, mn = max(Lr1, Lc1), min(Lr, Lc)
    cnt = int(p["count"])
    two = cnt > 1
    ro1, co1, ro2, co2 = (f32(v) for v in p["seed"])
    d1 = f32(co1 - ro1)
    f = np.zeros(16, np.float32)
    f[0] = _lg2(Lr1) / f32(16)
    f[1] = _lg2(Lc1) / f32(16)
    f[2] = mn / mx
    f[3] = _lg2(f32(1) + f32(cnt)) / f32(8)
    f[4] = f32(int(p["kscore"])) / f32(64)
    f[5] = d1 / mx
    f[6] = abs(f32(f32(co2 - ro2) - d1)) / mx if two else 0
    f[7] = ro1 / Lr1
    f[8] = co1 / Lc1
    f[9] = ro2 / Lr1 if two else 0
    f[10] = co2 / Lc1 if two else 0
    den = f32(mn - f32(5))
    f[11] = f32(cnt) / (den if den > 1 else f32(1))
    agree = int(np.sum((rmh == cmh) & (rmh != EMPTY_HASH)))
    f[12] = f32(agree) / f32(SKETCH)
    smin = f32(0)
    sdot = f32(0)
    for k in range(10):
        smin = f32(smin + min(rcomp[k], ccomp[k]))
        sdot = f32(sdot + f32(rcomp[k] * ccomp[k]))
    f[13], f[14] = smin, sdot
    a = min(ro1, co1)
    b = min(f32(Lr - ro1), f32(Lc - co1))
    f[15] = f32(a + b) / mx
    return f


def features(pairs, rseqs, cseqs):
    """Features of every pair; rseqs / cseqs are lists of code arrays."""
    rs = {}
    cs = {}
    out = np.zeros((len(pairs), 16), np.float32)
    for i, p in enumerate(pairs):
        r, c = int(p["r"]), int(p["c"])
        if r not in rs:
            rs[r] = seq_sketch(rseqs[r])
        if c not in cs:
            cs[c] = seq_sketch(cseqs[c])
        out[i] = pair_features(p, rs[r], cs[c])
    return out


def standardization(feats):
    """mean and 1/std per feature (float64 statistics, fp32 results)."""
    x = np.asarray(feats, np.float64)
    mu = x.mean(axis=0) if len(x) else np.zeros(16)
    sd = np.sqrt(((x - mu) ** 2).mean(axis=0)) if len(x) else np.zeros(16)
    istd = np.where(sd > 1e-6, 1.0 / np.where(sd > 0, sd, 1), 1.0)
    return mu.astype(np.float32), istd.astype(np.float32)


def standardize(feats, mean, istd):
    f = np.asarray(feats, np.float32)
    return ((f - mean).astype(np.float32) * istd).astype(np.float32)
