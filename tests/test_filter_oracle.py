"""Pins the CPU oracle of the learned prefilter (oracle/mlp_oracle.py).

bf16 rounding is checked against torch's bfloat16 conversion; the forward
pass, the gradient and Adam against torch float64 autograd / torch.optim on
the same bf16-rounded operands (straight-through rounding), so a dropped
term, a transposed weight or a wrong sign fails; the features against brute
force (set MinHash, direct composition counts) and closed forms.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M

from helpers import rand_seq

ALPHA = "ARNDCQEGHILKMFPSTWYVBZX*J"


def test_layout():
    assert M.NPARAM == 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 5313
    p = np.arange(M.NPARAM, dtype=np.float32)
    W1, b1, W2, b2, w3, b3 = M.unpack(p)
    assert W1.shape == (16, 64) and W1[1, 0] == 64 and b1[0] == 1024
    assert W2.shape == (64, 64) and W2[0, 1] == 1089 and b3[0] == 5312


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 20000).astype(np.float32),
                        np.array([0, -0.0, 1, 1 + 2 ** -8, 1 + 3 * 2 ** -9,
                                  65504, 1e-38, -3.3895e38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    np.testing.assert_array_equal(M.bf16(x), ref)


def _torch_forward(params, x, pos_weight, y):
    """float64 autograd model with straight-through bf16 rounding."""
    def ste(t):
        r = torch.from_numpy(M.bf16(t.detach().float().numpy())).double()
        return t + (r - t).detach()
    P = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    W1 = P[M.OFF_W1:M.OFF_B1].view(16, 64)
    b1 = P[M.OFF_B1:M.OFF_W2]
    W2 = P[M.OFF_W2:M.OFF_B2].view(64, 64)
    b2 = P[M.OFF_B2:M.OFF_W3]
    w3 = P[M.OFF_W3:M.OFF_B3]
    b3 = P[M.OFF_B3:]
    xb = torch.from_numpy(M.bf16(x)).double()
    h1 = ste(torch.relu(xb @ ste(W1) + b1))
    h2 = ste(torch.relu(h1 @ ste(W2) + b2))
    z3 = h2 @ ste(w3) + b3
    yt = torch.from_numpy(np.asarray(y, np.float64))
    p = torch.sigmoid(z3)
    loss = -(pos_weight * yt * torch.log(p) + (1 - yt) * torch.log(1 - p))
    loss = loss.mean()
    loss.backward()
    return z3.detach().numpy(), float(loss), P.grad.numpy()


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_and_grad_match_torch(seed):
    rng = np.random.default_rng(seed)
    params = M.init_params(seed)
    params[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.1, 64)
    params[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.1, 64)
    params[M.OFF_B3] = 0.3
    x = rng.normal(0, 1, (257, 16)).astype(np.float32)
    y = (rng.random(257) < 0.3).astype(np.float32)
    z3, _ = M.forward(params, x)
    loss, g = M.loss_and_grad(params, x, y, pos_weight=2.5)
    tz, tloss, tg = _torch_forward(params, x, 2.5, y)
    np.testing.assert_allclose(z3, tz, rtol=1e-4, atol=1e-4)
    assert abs(loss - tloss) < 1e-4 * max(1, abs(tloss))
    scale = np.abs(tg).max()
    np.testing.assert_allclose(g, tg, rtol=1e-3, atol=1e-5 * scale)
    # every block of the gradient is exercised
    for a, b in [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2),
                 (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)]:
        assert np.abs(g[a:b]).max() > 0


def test_grad_finite_difference_on_bias():
    """Biases enter unrounded, so the loss is differentiable in them."""
    rng = np.random.default_rng(3)
    params = M.init_params(3).astype(np.float64).astype(np.float32)
    x = rng.normal(0, 1, (64, 16)).astype(np.float32)
    y = (rng.random(64) < 0.5).astype(np.float32)
    _, g = M.loss_and_grad(params, x, y, 1.0)
    i = M.OFF_B3
    h = np.float32(1e-2)
    pp, pm = params.copy(), params.copy()
    pp[i] += h
    pm[i] -= h
    lp, _ = M.loss_and_grad(pp, x, y, 1.0)
    lm, _ = M.loss_and_grad(pm, x, y, 1.0)
    assert abs((lp - lm) / (2 * h) - g[i]) < 1e-3


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, M.NPARAM).astype(np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.tensor(p.copy(), dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 4):
        g = rng.normal(0, 1, M.NPARAM).astype(np.float32)
        M.adam_step(p, m, v, g, t, lr=1e-2)
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-6)


def _codes(s):
    return np.array([ALPHA.index(ch) for ch in s], np.uint8)


def test_murphy10_classes():
    # Murphy10ProteinAlph groups (inc/Alphabet.hpp:95-99)
    groups = ["DENQ", "FWY", "ILMV", "KR", "ST"]
    for g in groups:
        assert len({int(M.MURPHY10[ALPHA.index(ch)]) for ch in g}) == 1
    singles = "ACGHP"
    cls = [int(M.MURPHY10[ALPHA.index(ch)]) for ch in singles]
    cls += [int(M.MURPHY10[ALPHA.index(g[0])]) for g in groups]
    assert sorted(cls) == list(range(10))
    assert [int(M.MURPHY10[ALPHA.index(ch)]) for ch in "BZX*J"] == \
        [10, 11, 12, 13, 14]


def test_sketch_bruteforce():
    rng = np.random.default_rng(7)
    for L in [0, 3, 5, 6, 7, 40, 300]:
        s = rand_seq(rng, L)
        comp, mh, n = M.seq_sketch(s)
        assert n == L
        cls = [int(M.MURPHY10[c]) for c in s]
        std = [c for c in cls if c < 10]
        for k in range(10):
            want = np.float32(std.count(k)) / np.float32(max(len(std), 1))
            assert comp[k] == want
        kmers = set()
        for j in range(L - 5):
            code = 0
            for t in range(6):
                code = code * 15 + cls[j + t]
            kmers.add(code)
        for h in range(16):
            salt = (0x9E3779B9 * (h + 1)) & 0xFFFFFFFF
            if kmers:
                want = min(int(M.mix32(np.uint32(k ^ salt))) for k in kmers)
            else:
                want = 0xFFFFFFFF
            assert int(mh[h]) == want


def test_mix32_known_values():
    # lowbias32 constants; mix32(0) = 0 and the map is a bijection on a sample
    assert int(M.mix32(np.uint32(0))) == 0
    x = np.arange(1 << 16, dtype=np.uint32) * np.uint32(40503)
    assert len(np.unique(M.mix32(x))) == len(x)


def _pair(r, c, count, kscore, seed):
    p = np.zeros(1, [("r", "<u4"), ("c", "<u4"), ("count", "<u2"),
                     ("kscore", "<i2"), ("seed", "<u2", (4,))])
    p["r"], p["c"], p["count"], p["kscore"] = r, c, count, kscore
    p["seed"] = seed
    return p


def test_features_closed_forms():
    rng = np.random.default_rng(11)
    a = rand_seq(rng, 200)
    b = rand_seq(rng, 100)
    seqs = [a, b]
    # identical sequences, one seed on the main diagonal
    f = M.features(_pair(0, 0, 1, 30, [10, 10, 0, 0]), seqs, seqs)[0]
    assert f[2] == 1 and f[5] == 0 and f[6] == 0 and f[12] == 1
    assert abs(f[13] - 1) < 1e-6 and f[15] == 1
    assert f[0] == np.float32(np.log2(200)) / 16 and f[9] == 0 and f[10] == 0
    assert f[4] == np.float32(30 / 64)
    # two seeds, lengths 200 x 100
    f = M.features(_pair(0, 1, 3, -2, [50, 20, 120, 60]), seqs, seqs)[0]
    assert f[2] == np.float32(0.5)
    assert f[5] == np.float32(-30 / 200)
    assert f[6] == np.float32(abs((60 - 120) - (20 - 50)) / 200)
    assert f[3] == np.float32(np.log2(4) / 8) == np.float32(0.25)
    assert f[7] == np.float32(50 / 200) and f[8] == np.float32(20 / 100)
    assert f[9] == np.float32(120 / 200) and f[10] == np.float32(60 / 100)
    assert f[11] == np.float32(3 / 95)
    assert f[15] == np.float32((20 + min(150, 80)) / 200)
    # composition overlap of disjoint class sets is 0
    x = _codes("AAAAAAAAAA")
    y = _codes("GGGGGGGGGG")
    f = M.features(_pair(0, 1, 1, 0, [0, 0, 0, 0]), [x, y], [x, y])[0]
    assert f[13] == 0 and f[14] == 0 and f[12] == 0
    # empty sequences produce finite features
    e = np.zeros(0, np.uint8)
    f = M.features(_pair(0, 0, 0, 0, [0, 0, 0, 0]), [e], [e])[0]
    assert np.all(np.isfinite(f))


def test_standardization():
    rng = np.random.default_rng(2)
    x = rng.normal(3, 2, (1000, 16)).astype(np.float32)
    x[:, 5] = 7  # constant feature
    mu, istd = M.standardization(x)
    z = M.standardize(x, mu, istd)
    assert np.allclose(z.mean(0), 0, atol=1e-5)
    assert np.allclose(z[:, [i for i in range(16) if i != 5]].std(0), 1,
                       atol=1e-4)
    assert istd[5] == 1 and np.all(z[:, 5] == 0)
