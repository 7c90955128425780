"""The learned MC proposal: data collection, training and evaluation.

The network ranks the 16 candidate mutations of an MC state by their expected
energy after BFGS (see include/ud2.h, "learned proposal"). Everything that
touches the network's arithmetic runs in this project's HIP kernels:
candidates and their features come from the persistent MC kernel's collect
policy (ud2_lmc), training steps are ud2_mlp_train_step (forward, loss,
backward and Adam on MFMA), scores are ud2_mlp_forward. PyTorch only holds the
device buffers and shuffles group indices.

Training data: c1-style synthetic ligands from held-out seeds (9000 + i,
never used by benchmarks or tests), docked with the plain trajectory while
every step's 16 candidates are BFGS-refined to label them.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _native as N

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "proposal_weights.npz")
FEATURE_VERSION = 1  # bumps when the feature definition changes
HELD_OUT_SEED = 9000


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases, identity input normalisation."""
    rng = np.random.default_rng(seed)
    F, H = N.MLP_FEATURES, N.MLP_HIDDEN
    p = np.zeros(N.MLP_NPARAM + N.MLP_NNORM, np.float32)
    o = 0
    p[o:o + F * H] = rng.normal(0, math.sqrt(2.0 / F), F * H); o += F * H
    o += H
    p[o:o + H * H] = rng.normal(0, math.sqrt(2.0 / H), H * H); o += H * H
    o += H
    p[o:o + H] = rng.normal(0, math.sqrt(1.0 / H), H); o += H
    p[N.MLP_NPARAM + F:] = 1.0
    return p


def bf16_to_float(x):
    """bf16 bit patterns (int16 tensor) -> float32 tensor (same device)."""
    import torch
    return (x.to(torch.int32) << 16).view(torch.float32)


@dataclass
class Dataset:
    x: object        # int16 device tensor [groups * 16, 80], bf16 bits
    y: object        # float32 device tensor [groups * 16], energies after BFGS
    n_groups: int
    e0: object = None  # float32 device tensor [groups * 16], energies before BFGS (optional)


def collect(engine, ligands, nchain: int = 64, steps: int = 8, opt_steps: int = -1, seed: int = 1,
            lig_index_base: int = HELD_OUT_SEED, with_e0: bool = False, stream=None) -> Dataset:
    """Runs the collect policy of the learned MC kernel over the ligands."""
    import torch
    from .engine import heuristic_opt_steps
    ls = N.LigSet(ligands)
    engine.ensure_types(ligands)
    pb = N.PoseBuffers(ls, nchain)
    pb.init()
    npose = ls.n_lig * nchain
    x = torch.zeros((steps * npose * N.MLP_GROUP, N.MLP_FEATURES), dtype=torch.int16, device="cuda")
    y = torch.zeros(steps * npose * N.MLP_GROUP, dtype=torch.float32, device="cuda")
    e0 = torch.zeros_like(y) if with_e0 else None
    if opt_steps < 0:
        opt_steps = heuristic_opt_steps(max(l.natom for l in ligands))
    prm = N.LmcParams()
    prm.mc = N.McParams(mc_steps=steps, opt_steps=opt_steps, randomize=1, seed=seed & 0xFFFFFFFF,
                        mode=N.MODE_GRID, lig_index_base=lig_index_base, stats_d=None)
    prm.policy, prm.n_cand, prm.top_k = N.POLICY_COLLECT, N.MLP_GROUP, 1
    prm.collect_x_d, prm.collect_y_d = x.data_ptr(), y.data_ptr()
    prm.collect_e0_d = e0.data_ptr() if with_e0 else None
    N.check(N.lib().ud2_lmc(engine.scene.h, ls.h, pb.ref, C.byref(prm), None, N.stream_ptr(stream)), "ud2_lmc")
    torch.cuda.synchronize()
    return Dataset(x, y, steps * npose, e0)


def finite_groups(data: Dataset) -> np.ndarray:
    """Groups whose 16 feature rows are all finite and that have a finite label."""
    import torch
    xf = bf16_to_float(data.x).view(data.n_groups, N.MLP_GROUP * N.MLP_FEATURES)
    ok = torch.isfinite(xf).all(dim=1) & torch.isfinite(data.y.view(data.n_groups, N.MLP_GROUP)).any(dim=1)
    return torch.nonzero(ok).view(-1).cpu().numpy().astype(np.int32)


def fit_normalization(params: np.ndarray, data: Dataset, groups: np.ndarray, sample_groups: int = 65536,
                      seed: int = 0) -> None:
    """Sets mu / inv_sigma from a sample of the training groups' inputs (in place)."""
    import torch
    g = torch.as_tensor(np.random.default_rng(seed).choice(groups, min(sample_groups, len(groups)),
                                                           replace=False), device="cuda")
    rows = (g[:, None] * N.MLP_GROUP + torch.arange(N.MLP_GROUP, device="cuda")[None, :]).reshape(-1)
    xs = bf16_to_float(data.x[rows])
    mu = xs.mean(0)
    sd = xs.std(0)
    inv = torch.where(sd > 1e-6, 1.0 / sd, torch.ones_like(sd))
    F = N.MLP_FEATURES
    params[N.MLP_NPARAM:N.MLP_NPARAM + F] = mu.cpu().numpy()
    params[N.MLP_NPARAM + F:] = inv.cpu().numpy()


def split_groups(groups: np.ndarray, val_frac: float = 0.05, seed: int = 0):
    perm = np.random.default_rng(seed).permutation(groups).astype(np.int32)
    nv = max(1, int(len(perm) * val_frac))
    return perm[nv:], perm[:nv]


def scores(mlp: N.Mlp, data: Dataset, groups: np.ndarray, stream=None) -> np.ndarray:
    """Network scores [len(groups), 16] through ud2_mlp_forward."""
    import torch
    g = torch.as_tensor(np.ascontiguousarray(groups, np.int32), device="cuda")
    y = torch.zeros(len(groups) * N.MLP_GROUP, device="cuda")
    N.check(N.lib().ud2_mlp_forward(mlp.h, C.c_void_p(data.x.data_ptr()), C.c_void_p(g.data_ptr()), len(groups),
                                    C.c_void_p(y.data_ptr()), N.stream_ptr(stream)), "ud2_mlp_forward")
    torch.cuda.synchronize()
    return y.cpu().numpy().reshape(len(groups), N.MLP_GROUP)


def ranking_quality(pred: np.ndarray, label: np.ndarray, top_k: int = 4) -> dict:
    """How well the scores order the candidates of each group, against random order.

    best_in_topk: fraction of groups whose best candidate (lowest label) is in
    the predicted top_k (random: top_k / 16). regret_top1: mean label of the
    predicted best minus the true best (random: mean label minus best).
    """
    lab = np.where(np.isfinite(label), label, np.inf)
    ok = np.isfinite(lab).any(axis=1)
    pred, lab = pred[ok], lab[ok]
    lab = np.minimum(lab, lab.min(axis=1, keepdims=True) + 6.0)  # the clipped labels the loss sees
    best = lab.argmin(axis=1)
    order = np.argsort(pred, axis=1, kind="stable")
    in_top = (order[:, :top_k] == best[:, None]).any(axis=1)
    top1 = lab[np.arange(len(lab)), order[:, 0]]
    topk_best = np.take_along_axis(lab, order[:, :top_k], axis=1).min(axis=1)
    return dict(groups=int(len(lab)), best_in_topk=float(in_top.mean()), random_best_in_topk=top_k / 16.0,
                regret_top1=float(np.mean(top1 - lab.min(axis=1))),
                random_regret_top1=float(np.mean(lab.mean(axis=1) - lab.min(axis=1))),
                regret_topk_best=float(np.mean(topk_best - lab.min(axis=1))))


def train(mlp: N.Mlp, data: Dataset, train_groups: np.ndarray, epochs: int = 4, batch_groups: int = 1024,
          lr: float = 1e-3, seed: int = 0, log=print, stream=None) -> list:
    """Adam over shuffled minibatches of groups with ud2_mlp_train_step; returns the loss per epoch."""
    import torch
    rng = np.random.default_rng(seed)
    opt = N.Adam(lr, 0.9, 0.999, 1e-8)
    loss = torch.zeros(1, device="cuda")
    hist = []
    st = N.stream_ptr(stream)
    for ep in range(epochs):
        perm = torch.as_tensor(rng.permutation(train_groups).astype(np.int32), device="cuda")
        nb = len(perm) // batch_groups
        t0 = time.time()
        acc = torch.zeros(1, device="cuda")
        for b in range(nb):
            g = perm[b * batch_groups:(b + 1) * batch_groups]
            N.check(N.lib().ud2_mlp_train_step(mlp.h, C.c_void_p(data.x.data_ptr()), C.c_void_p(data.y.data_ptr()),
                                               C.c_void_p(g.data_ptr()), batch_groups, C.byref(opt),
                                               C.c_void_p(loss.data_ptr()), st), "ud2_mlp_train_step")
            acc += loss
        torch.cuda.synchronize()
        hist.append(float(acc.item()) / max(nb, 1))
        log(f"epoch {ep}: loss {hist[-1]:.4f} ({nb} steps, {time.time() - t0:.1f} s)")
    return hist


def save(params: np.ndarray, path: str = WEIGHTS, **meta) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=np.asarray(params, np.float32), feature_version=np.int32(FEATURE_VERSION),
                        **{k: np.asarray(v) for k, v in meta.items()})


def load(path: str = WEIGHTS) -> np.ndarray:
    """Parameters of a saved network (numeric arrays only, no pickles)."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["feature_version"]) != FEATURE_VERSION:
            raise ValueError(f"{path}: feature version {int(z['feature_version'])}, expected {FEATURE_VERSION}")
        p = np.asarray(z["params"], np.float32)
    if p.size != N.MLP_NPARAM + N.MLP_NNORM:
        raise ValueError(f"{path}: {p.size} parameters, expected {N.MLP_NPARAM + N.MLP_NNORM}")
    return p


def build_training_set(engine, n_lig: int = 512, nchain: int = 128, steps: int = 10, seed: int = 1, log=print,
                       with_e0: bool = False):
    """Collects the held-out training set (c1-style ligands, seeds 9000 + i)."""
    from . import synthetic as S
    ligs = [s.ligand for s in S.config1_ligands(n_lig, base_seed=HELD_OUT_SEED)]
    t0 = time.time()
    data = collect(engine, ligs, nchain=nchain, steps=steps, seed=seed, with_e0=with_e0)
    log(f"collected {data.n_groups} groups ({data.n_groups * N.MLP_GROUP} candidates) in {time.time() - t0:.1f} s")
    return data


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Train the learned MC proposal network on the GPU.")
    ap.add_argument("--n-lig", type=int, default=512)
    ap.add_argument("--nchain", type=int, default=128)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch-groups", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--dump-groups", type=int, default=0, help="also save this many random groups (x, y, e0)")
    ap.add_argument("--dump", default="")
    a = ap.parse_args(argv)
    from . import synthetic as S
    from .engine import DockEngine
    eng = DockEngine(S.make_receptor(), S.docking_box(), mode=N.MODE_GRID)
    data = build_training_set(eng, a.n_lig, a.nchain, a.steps, seed=a.seed + 1, log=lambda s: print(s, flush=True),
                              with_e0=a.dump_groups > 0)
    if a.dump_groups > 0:
        import torch
        g = np.sort(np.random.default_rng(a.seed + 7).choice(data.n_groups, a.dump_groups, replace=False))
        rows = (g[:, None] * N.MLP_GROUP + np.arange(N.MLP_GROUP)[None, :]).reshape(-1)
        r = torch.as_tensor(rows, device="cuda")
        np.savez_compressed(a.dump, x=data.x[r].cpu().numpy(), y=data.y[r].cpu().numpy(),
                            e0=data.e0[r].cpu().numpy(), groups=g)
        print("dumped", a.dump, flush=True)
    good = finite_groups(data)
    print(f"{len(good)} of {data.n_groups} groups have finite features and labels", flush=True)
    tr, va = split_groups(good, seed=a.seed)
    params = init_params(a.seed)
    fit_normalization(params, data, tr, seed=a.seed)
    mlp = N.Mlp(params)
    before = ranking_quality(scores(mlp, data, va), data.y.view(-1, N.MLP_GROUP)[va].cpu().numpy())
    print("untrained:", before, flush=True)
    hist = train(mlp, data, tr, a.epochs, a.batch_groups, a.lr, a.seed, log=lambda s: print(s, flush=True))
    after = ranking_quality(scores(mlp, data, va), data.y.view(-1, N.MLP_GROUP)[va].cpu().numpy())
    print("trained:", after, flush=True)
    save(mlp.params(), a.out, loss=np.asarray(hist, np.float32), n_groups=data.n_groups,
         best_in_topk=after["best_in_topk"], regret_top1=after["regret_top1"])
    print("saved", a.out, flush=True)


if __name__ == "__main__":
    main()
