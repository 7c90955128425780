"""The learned proposal's oracle (features, network forward / backward, loss,
Adam) pinned by independent restatements (torch's bfloat16 cast, torch
matrix products and autograd in float64, torch.optim.Adam), finite
differences and hand-computed feature values on affine maps. CPU only."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O

NP, NF, NH, NN, G = O.mlp_sizes()


def rand_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP + NN, np.float32)
    p[:NF * NH] = rng.normal(0, scale * math.sqrt(2.0 / NF), NF * NH)
    p[NF * NH:NF * NH + NH] = rng.normal(0, 0.1, NH)
    o = NF * NH + NH
    p[o:o + NH * NH] = rng.normal(0, scale * math.sqrt(2.0 / NH), NH * NH)
    o += NH * NH
    p[o:o + NH] = rng.normal(0, 0.1, NH)
    o += NH
    p[o:o + NH] = rng.normal(0, math.sqrt(1.0 / NH), NH)
    p[NP - 1] = 0.3
    p[NP:NP + NF] = rng.normal(0, 0.5, NF)            # mu
    p[NP + NF:] = rng.uniform(0.5, 2.0, NF)            # inv_sigma
    return p


def bits_to_f32(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    t = torch.as_tensor(p[:NP], dtype=torch.float64)
    o = 0
    W1 = t[o:o + NF * NH].reshape(NF, NH); o += NF * NH
    b1 = t[o:o + NH]; o += NH
    W2 = t[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = t[o:o + NH]; o += NH
    w3 = t[o:o + NH]; o += NH
    b3 = t[o]
    return W1, b1, W2, b2, w3, b3


def bf16(t):
    return t.to(torch.float32).to(torch.bfloat16).to(torch.float64)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 3.4e38, -3.4e38, np.inf, -np.inf]])
    x = x.astype(np.float32)
    ref = torch.as_tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = O.bf16_round(x)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert O.bf16_round(np.float32([1.0 + 2.0 ** -8]))[0] == 1.0                # tie to even
    assert O.bf16_round(np.float32([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    assert np.isnan(O.bf16_round(np.float32([np.nan]))[0])
    assert np.array_equal(bits_to_f32(O.bf16_bits(x)).view(np.uint32), ref.view(np.uint32))


def test_forward_matches_torch_restatement():
    p = rand_params(2)
    rng = np.random.default_rng(3)
    x = O.bf16_bits(rng.normal(0, 2, (64, NF)).astype(np.float32))
    y = O.mlp_forward(p, x)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    mu = torch.as_tensor(p[NP:NP + NF]); inv = torch.as_tensor(p[NP + NF:])
    xf = torch.as_tensor(bits_to_f32(x))
    xn = ((xf - mu) * inv).to(torch.bfloat16).to(torch.float64)
    a1 = bf16(torch.relu(xn @ bf16(W1) + b1))
    yt = torch.relu(a1 @ bf16(W2) + b2) @ bf16(w3) + b3
    assert np.allclose(y, yt.numpy(), rtol=1e-10, atol=1e-10)
    # the exact network differs from the rounded one by bf16-sized amounts only
    ye = O.mlp_forward(p, x, round_=False)
    assert np.abs(ye - y).max() < 0.05 * (1 + np.abs(y).max())


def test_backward_exact_matches_autograd():
    p = rand_params(4)
    rng = np.random.default_rng(5)
    xb = O.bf16_bits(rng.normal(0, 2, (32, NF)).astype(np.float32))
    dy = rng.normal(0, 1, 32)
    g = O.mlp_backward(p, xb, dy, round_=False)
    W1, b1, W2, b2, w3, b3 = [t.clone().requires_grad_(True) for t in unpack(p)]
    mu = torch.as_tensor(p[NP:NP + NF], dtype=torch.float64); inv = torch.as_tensor(p[NP + NF:], dtype=torch.float64)
    xf = torch.as_tensor(bits_to_f32(xb), dtype=torch.float64)
    xn = ((xf.float() - mu.float()) * inv.float()).double()
    y = torch.relu(torch.relu(xn @ W1 + b1) @ W2 + b2) @ w3 + b3
    (torch.as_tensor(dy) * y).sum().backward()
    ref = torch.cat([W1.grad.reshape(-1), b1.grad, W2.grad.reshape(-1), b2.grad, w3.grad, b3.grad.reshape(1)])
    assert np.allclose(g, ref.numpy(), rtol=1e-9, atol=1e-11)


def test_backward_straight_through_matches_autograd():
    """With rounding on, the oracle differentiates straight through every bf16 rounding."""
    p = rand_params(6)
    rng = np.random.default_rng(7)
    xb = O.bf16_bits(rng.normal(0, 2, (16, NF)).astype(np.float32))
    dy = rng.normal(0, 1, 16)
    g = O.mlp_backward(p, xb, dy, round_=True)
    W1, b1, W2, b2, w3, b3 = [t.clone().requires_grad_(True) for t in unpack(p)]
    ste = lambda t: t + (bf16(t) - t).detach()  # noqa: E731
    mu = torch.as_tensor(p[NP:NP + NF]); inv = torch.as_tensor(p[NP + NF:])
    xn = ((torch.as_tensor(bits_to_f32(xb)) - mu) * inv).to(torch.bfloat16).to(torch.float64)
    a1 = ste(torch.relu(xn @ ste(W1) + b1))
    y = torch.relu(a1 @ ste(W2) + b2) @ ste(w3) + b3
    (torch.as_tensor(dy) * y).sum().backward()
    ref = torch.cat([W1.grad.reshape(-1), b1.grad, W2.grad.reshape(-1), b2.grad, w3.grad, b3.grad.reshape(1)])
    assert np.allclose(g, ref.numpy(), rtol=1e-9, atol=1e-11)


def test_backward_finite_differences():
    p = rand_params(8)
    rng = np.random.default_rng(9)
    x = rng.normal(0, 2, (8, NF))
    xb = O.bf16_bits(x.astype(np.float32))
    xin = bits_to_f32(xb).astype(np.float64)
    dy = rng.normal(0, 1, 8)
    g = O.mlp_backward(p, xb, dy, round_=False)
    idx = np.r_[rng.integers(0, NF * NH, 6), NF * NH + rng.integers(0, NH, 3),
                NF * NH + NH + rng.integers(0, NH * NH, 6), NP - 2 * NH - 1 + rng.integers(0, NH, 2),
                NP - NH - 1 + rng.integers(0, NH, 3), NP - 1]
    # the exact network is fed unrounded normalised inputs computed like the oracle's float path
    mu, inv = p[NP:NP + NF], p[NP + NF:]
    xn_in = ((xin.astype(np.float32) - mu) * inv).astype(np.float64)
    pz = p.copy()
    pz[NP:NP + NF] = 0.0
    pz[NP + NF:] = 1.0
    for i in idx:
        h = 1e-3 * max(1.0, abs(float(p[i])))
        pp, pm = pz.copy(), pz.copy()
        pp[i] += h
        pm[i] -= h
        step = float(pp[i]) - float(pm[i])
        fd = (dy @ O.mlp_forward_exact(pp, xn_in) - dy @ O.mlp_forward_exact(pm, xn_in)) / step
        assert abs(fd - g[i]) < 1e-5 * (1 + abs(g[i])), (i, fd, g[i])


def test_group_loss_closed_forms():
    rng = np.random.default_rng(10)
    # yhat = label + a constant per group: zero loss and gradient
    lab = rng.normal(-5, 1, (4, G)).astype(np.float32).clip(-8, -3)
    yh = lab.astype(np.float64) + rng.normal(0, 3, (4, 1))
    loss, d = O.group_loss(yh.reshape(-1), lab.reshape(-1))
    assert abs(loss) < 1e-12 and np.abs(d).max() < 1e-12
    # hand case: yhat = e_0, labels 0
    yh = np.zeros(G); yh[0] = 1.0
    loss, d = O.group_loss(yh, np.zeros(G, np.float32))
    assert abs(loss - 0.05859375) < 1e-15
    assert abs(d[0] - 2 * (15 / 16) / 16) < 1e-15 and abs(d[1] + 2 * (1 / 16) / 16) < 1e-15
    # clipping at min + 6: labels above the clip are interchangeable, nan / inf included
    lab = rng.normal(-6, 1, G).astype(np.float32)
    lab[3] = 1e6
    yh = rng.normal(-6, 1, G)
    l1, _ = O.group_loss(yh, lab)
    lab2 = lab.copy(); lab2[3] = np.inf
    lab3 = lab.copy(); lab3[3] = lab.min() + 7.0
    assert abs(O.group_loss(yh, lab2)[0] - l1) < 1e-12 and abs(O.group_loss(yh, lab3)[0] - l1) < 1e-12
    # a group without a finite label contributes nothing
    loss, d = O.group_loss(rng.normal(0, 1, G), np.full(G, np.nan, np.float32))
    assert loss == 0.0 and np.all(d == 0.0)
    # gradient = finite differences of the loss
    yh = rng.normal(-6, 2, 2 * G)
    lab = rng.normal(-6, 3, 2 * G).astype(np.float32)
    _, d = O.group_loss(yh, lab)
    for i in (0, 5, 17, 31):
        e = np.zeros_like(yh); e[i] = 1e-6
        fd = (O.group_loss(yh + e, lab)[0] - O.group_loss(yh - e, lab)[0]) / 2e-6
        assert abs(fd - d[i]) < 1e-8


def test_adam_matches_torch():
    rng = np.random.default_rng(11)
    p = rng.normal(0, 1, 257).astype(np.float32)
    m = np.zeros_like(p); v = np.zeros_like(p)
    tp = torch.nn.Parameter(torch.as_tensor(p.copy()))
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 4):
        g = rng.normal(0, 1, 257).astype(np.float32)
        O.adam_step(p, m, v, g, 1e-3, 0.9, 0.999, 1e-8, step)
        tp.grad = torch.as_tensor(g)
        opt.step()
        assert np.allclose(p, tp.detach().numpy(), rtol=1e-5, atol=1e-7)


def test_features_hand_computed_on_affine_maps():
    """Maps affine in x make the trilinear value and gradient exact; every statistic of
    every feature class and every move global is checked against hand values."""
    from unidock_engine import topology as T
    n = (11, 11, 11)
    xs = np.arange(11, dtype=np.float32)
    lin = lambda a, b: np.broadcast_to(a + b * xs[None, None, :], (11, 11, 11))  # noqa: E731
    maps = np.stack([lin(-1.0, 0.1), lin(0.0, 2.0), lin(-0.2, 0.0), lin(-100.0, 20.0)]).astype(np.float32)
    slot = [-1] * 21
    slot[2], slot[3], slot[6], slot[10] = 0, 1, 2, 3
    rec = T.Receptor(xyz=np.zeros((1, 3), np.float32), types=np.asarray([2], np.int32))
    box = T.Box(lo=np.zeros(3, np.float32), hi=np.full(3, 10.0, np.float32))
    sc = O.Scene(rec, box, "grid", maps=maps, grid_n=n, grid_origin=box.lo, grid_spacing=1.0, slot_of_type=slot)
    coords = [[2, 5, 5], [3, 5, 5], [4, 5, 5], [5, 5, 12], [8, 5, 5], [4, 5, 5]]
    types = [2, 3, 0, 6, 2, 10]
    lig = T.make_ligand("k", np.asarray(coords, np.float32), types, [], [], [], [], np.zeros((0, 2), np.int32))
    f = O.features(lig, sc, np.asarray(coords, np.float64).reshape(-1), kind=2, magnitude=0.25, e_acc=-35.0,
                   moved=0.5)
    exp = np.zeros(80)
    exp[0:8] = [2, -1.0, 0, 0, 1, 0, 0.2, 0]                      # C_H: e = -0.8, -0.2; |g| = 0.1
    exp[8:16] = [1, 0, 1, math.log(7.0), 0, 0, 0, 2]              # C_P: e = 6, |g| = 2
    exp[32:40] = [1, -0.2, 0, 0, 0, 2, 0, 0]                      # N_A: z clamped, 2 A outside the box
    exp[40:48] = [1, -20, 0, 0, 1, 0, 8, 0]                       # O_A: e = -20, |g| = 20 capped at 8
    curl = -0.8 + 6.0 / 7.0 - 0.2 - 0.2 - 20.0
    exp[64:73] = [0, 0, 1, 0.25, 5 / 64, 0, -2.0, 0.5, max(curl, -20.0) / 10]
    assert np.allclose(f, exp, atol=1e-6), np.nonzero(~np.isclose(f, exp, atol=1e-6))
    g = O.features(lig, sc, np.asarray(coords, np.float64).reshape(-1))
    assert np.all(g[64:67] == 0) and g[67] == 0 and g[70] == 0 and g[71] == 0
