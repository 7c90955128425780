#!/usr/bin/env python
"""Docking throughput on one MI355X: ligands/sec over the 1,024-ligand uneven batch,
learned proposal vs plain MC.

Workload (synthetic, seeded): receptor of 2,000 heavy atoms in a shell around
the pocket, 30 A box, FP32 grid maps for 8 atom types at 0.375 A (81^3 points);
1,024 ligands with 10-60 heavy atoms and 0-15 rotatable bonds. One "step" = one
docking pass over the whole batch with the inputs already resident on the GPU:
MC search, clustering at 1.0 A, refinement (5 rounds x 2 BFGS iterations with
the penalty-slope ramp), filtering (energy_range 10, num_pose 10) and final
scoring.

  plain    128 chains x 40 MC steps, one random mutation + BFGS of
           (25 + natom_max) / 3 iterations per step (the reference's search);
  learned  the learned engine: 36 chains x 40 MC steps; each chain starts from
           the best of 4 randomised starts by the network, and each step draws
           16 mutations, scores them with the network on MFMA and samples one
           with p ~ exp(-score / 1 kcal/mol) for BFGS and the Metropolis test.

The plain engine is also timed at the learned engine's budget (36 x 40,
``config.plain_same_budget``): on this workload that control shows how much
of the speedup comes from the smaller budget rather than from the network.

``value`` is the learned engine's ligands/s; the plain engine is timed on the
same batch in the same run (``config.plain``) with the median top-pose
affinities of both. ``roofline`` covers the learned MC kernel: achieved =
grid-gather bytes of one launch (energy evaluations and candidate features x
32 B x heavy atoms, counted by the kernel) / its average duration measured
with HIP events on its stream. ``cpu_baseline`` times the CPU oracle's whole
per-ligand plain pipeline (direct receptor sum, as the reference scores) on a
bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "uni-dock2_amd"))

METRIC = "ligands/sec @1,024 uneven batch, learned vs plain, top score within 0.1 kcal/mol"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-lig", type=int, default=1024)
    ap.add_argument("--nchain", type=int, default=128)
    ap.add_argument("--mc-steps", type=int, default=40)
    ap.add_argument("--mode", choices=["grid", "direct"], default="grid")
    ap.add_argument("--cpu-sample", type=int, default=2, help="ligands timed on the CPU oracle (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["batch", "c4", "score", "copy", "gather"], default="batch",
                    help="batch: docking throughput (headline); c4: the large uneven batch (4,096 ligands, 10-80 "
                         "heavy atoms, 0-20 rotatable bonds, 256 chains x 40 steps for both engines, one timed pass "
                         "each after a 16-ligand warmup); score: E+gradient microbenchmark; copy: a 256 MiB device "
                         "copy that calibrates the FETCH_SIZE / WRITE_SIZE counters; gather: random 32-byte map "
                         "gathers (the gather ceiling)")
    ap.add_argument("--search", choices=["learned", "plain"], default="learned",
                    help="learned: headline (learned vs plain); plain: the plain engine only")
    ap.add_argument("--learned-nchain", type=int, default=36)
    ap.add_argument("--learned-steps", type=int, default=40)
    ap.add_argument("--n-cand", type=int, default=16)
    ap.add_argument("--top-k", type=int, default=1)
    ap.add_argument("--weights", default="", help="network parameters (.npz); default: the packaged ones")
    ap.add_argument("--sel-temp", type=float, default=1.0,
                    help="learned: sample one candidate with p ~ exp(-score / T) (0: top-k exact rescoring)")
    ap.add_argument("--n-start", type=int, default=4, help="learned: randomised starts per chain, best by the network")
    ap.add_argument("--no-control", action="store_true",
                    help="learned: skip the plain engine at the learned engine's chain / step budget")
    ap.add_argument("--n-pose", type=int, default=1 << 20)
    return ap.parse_args()


def host_cpu() -> dict:
    """The host's CPU model and the cores this process may use (lscpu / nproc)."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}


def cpu_baseline(ligands, receptor, box, nchain, mc_steps, opt_steps, n_sample, threads):
    """CPU oracle, whole DockTask pipeline per ligand, direct receptor sum, FP32 arithmetic."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    O.build()
    sc = O.Scene(receptor, box, "direct")
    secs = 0.0
    evals = 0
    for i, lig in enumerate(ligands[:n_sample]):
        topo = O.Topo(lig)
        start = O.PoseArr.initial(lig, dtype=__import__("numpy").float32)
        _, _, dt, ev = O.dock_ligand(topo, sc, start, i, nchain, mc_steps, opt_steps, 5, 12345, 1.0, 10.0, 10,
                                     True, nthreads=threads, sfx="f32")
        secs += dt
        evals += ev
    return {
        "value": round(n_sample / secs, 6),
        "unit": "ligands/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"first {n_sample} ligands of the batch, full pipeline ({nchain} chains x {mc_steps} MC steps, "
                   f"BFGS {opt_steps}, cluster, refine, score), direct receptor sum, FP32, {threads} threads, "
                   f"{secs:.1f} s, {evals} energy evaluations"),
        "host": host_cpu(),
    }


def score_microbench(args):
    """E + gradient of 1,048,576 random poses of the 40-atom, 8-torsion ligand (grid inter term,
    curl cap 10): evals/s and the HBM roofline of 1,400 algorithmic bytes per eval (32 B of map
    corners per heavy atom, 60 B of pose vector in, 60 B of energy and gradient out)."""
    import numpy as np
    import torch
    from unidock_engine import _native as N
    from unidock_engine import synthetic as S

    N.check(N.lib().ud2_set_device(0), "ud2_set_device")
    receptor = S.make_receptor()
    box = S.docking_box()
    lig = S.config2_ligand().ligand
    mode = N.MODE_GRID if args.mode == "grid" else N.MODE_DIRECT
    scene = N.Scene(receptor, box)
    if mode == N.MODE_GRID:
        scene.build_grids(S.GRID_SPACING, [2, 3, 4, 5, 6, 10, 11, 15])
    ls = N.LigSet([lig])
    n = args.n_pose
    pv = torch.as_tensor(S.random_pose_vectors(n, lig.ntors, box, seed=3), device="cuda")
    e = torch.zeros(n, device="cuda")
    g = torch.zeros((n, 6 + lig.ntors), device="cuda")
    st = N.stream_ptr()

    def run():
        N.check(N.lib().ud2_score_posevecs(scene.h, ls.h, 0, n, C.c_void_p(pv.data_ptr()),
                                           C.c_void_p(e.data_ptr()), C.c_void_p(g.data_ptr()), mode, 10.0, 1e6, st),
                "ud2_score_posevecs")

    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    evs = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run()
        b.record()
        evs.append((a, b))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    ms = np.asarray([a.elapsed_time(b) for a, b in evs])
    # algorithmic bytes per eval: 8 map corners per heavy atom, the pose vector in (centre, quaternion,
    # torsions) and the energy and gradient out
    bytes_per_eval = 32 * lig.n_heavy + 4 * (7 + lig.ntors) + 4 * (1 + 6 + lig.ntors)
    achieved = n * bytes_per_eval / (ms.mean() / 1000.0) / 1e9
    out = {
        "metric": "energy+gradient evals/sec (scoring microbenchmark)",
        "value": round(n * args.steps / wall, 1),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded random poses)",
        "config": {"workload": f"{n} random poses of a {lig.natom}-atom, {lig.ntors}-torsion ligand, "
                               f"{lig.npair} intra pairs, {args.mode} inter term",
                   "kernel_ms": round(float(ms.mean()), 3)},
        "roofline": {"kernel": "posevec_kernel<1>", "bound": "hbm", "achieved": round(achieved, 3),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": None, "algorithmic_bytes_per_launch": int(n * bytes_per_eval)},
        "cpu_baseline": None,
    }
    if n == 1 << 20 and args.mode == "grid":
        t = profiled_traffic("posevec_traffic_bytes_per_launch", SCORE_TRAFFIC_SUMMARY, "1048576 poses score grid")
        if t:
            out["roofline"]["traffic"] = t
            out["roofline"]["traffic_source"] = "profiles/posevec_kernel_c2_score.md (FETCH_SIZE + WRITE_SIZE)"
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        m = min(n, 1 << 16)
        osc = O.Scene(receptor, box, "direct")
        _, _, dt = O.score_posevecs(O.Topo(lig), osc, pv[:m].cpu().numpy(), v_cap=10.0, slope=1e6,
                                    nthreads=threads, sfx="f32")
        out["cpu_baseline"] = {"value": round(m / dt, 1), "unit": "evals/s", "cores": threads, "kind": "port",
                               "sample": f"{m} of the poses, direct receptor sum, FP32, {threads} threads, {dt:.1f} s",
                               "host": host_cpu()}
    print(json.dumps(out), flush=True)


def gather_probe(scene, span_cells: int = 0, n_gathers: int = 1 << 27, reps: int = 5) -> dict:
    """Random 32-byte gathers over the scene's cell-major maps (ud2_scene_gather_probe): the
    bandwidth scattered map reads can reach over that footprint, timed with HIP events."""
    import numpy as np
    import torch
    from unidock_engine import _native as N
    n_thr = (n_gathers + 15) // 16
    out = torch.empty(n_thr, dtype=torch.float32, device="cuda")
    st = N.stream_ptr()

    def run(seed):
        N.check(N.lib().ud2_scene_gather_probe(scene.h, n_gathers, span_cells, seed, C.c_void_p(out.data_ptr()), st),
                "ud2_scene_gather_probe")

    run(0)
    ms = []
    for r in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        run(r + 1)
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    gn, _, _, n_types = scene.grid_info()
    total = int(gn[0] - 1) * int(gn[1] - 1) * int(gn[2] - 1) * int(n_types)
    span = span_cells if span_cells > 0 else total
    return {"achieved": round(n_gathers * 32 / (float(np.median(ms)) / 1e3) / 1e9, 1), "unit": "GB/s",
            "footprint_bytes": int(span * 32), "gathers": n_gathers, "ms": round(float(np.median(ms)), 3)}


def gather_microbench(args):
    """Random-gather ceiling of the grid maps at three footprints: all maps (the search's), one map,
    and 4 MB (one XCD's L2)."""
    from unidock_engine import _native as N
    from unidock_engine import synthetic as S
    N.check(N.lib().ud2_set_device(0), "ud2_set_device")
    scene = N.Scene(S.make_receptor(), S.docking_box())
    scene.build_grids(S.GRID_SPACING, [2, 3, 4, 5, 6, 10, 11, 15])
    gn, _, _, n_types = scene.grid_info()
    per_map = int(gn[0] - 1) * int(gn[1] - 1) * int(gn[2] - 1)
    rows = {name: gather_probe(scene, span) for name, span in
            (("all_maps", 0), ("one_map", per_map), ("l2_4MB", (4 << 20) // 32))}
    full = rows["all_maps"]
    print(json.dumps({
        "metric": "random 32-byte map gathers, GB/s", "value": full["achieved"], "unit": "GB/s", "n_gpus": 1,
        "steps": 5, "warmup": 1, "ms_per_step": full["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic receptor grid maps",
        "config": {"workload": f"{full['gathers']} random cell gathers (two float4 each) per launch", **rows},
        "roofline": {"bound": "hbm", "achieved": full["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(full["achieved"] / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None}), flush=True)


def copy_probe():
    """A plain 256 MiB device-to-device copy, three times (counter calibration)."""
    import torch
    n = 256 << 20
    a = torch.empty(n // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_per_copy": n, "copies": 3}), flush=True)


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "lmc_mc_kernels_c1_batch_traffic.md")
SCORE_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "posevec_kernel_c2_score.md")


DEFAULT_TRAFFIC_CONFIG = "1024 ligands grid; learned 36x40 starts 4 cand 16 T 1.0; plain 128x40"


def kernel_sources_digest() -> str:
    """SHA-1 (12 hex digits) of the code that defines and launches the kernels (the device sources and
    ud2_api.cpp, comments and blank lines left out; not the whole-run orchestration): a traffic summary
    counts only for the build it measured."""
    import hashlib
    import re
    d = os.path.join(ROOT, "uni-dock2_amd", "csrc")
    h = hashlib.sha1()
    for f in sorted(os.listdir(d)):
        if f.endswith((".h", ".hip")) or f == "ud2_api.cpp":
            text = open(os.path.join(d, f), encoding="utf-8").read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            text = re.sub(r"//[^\n]*", "", text)
            code = "\n".join(line.rstrip() for line in text.splitlines() if line.strip())
            h.update(f.encode())
            h.update(code.encode())
    return h.hexdigest()[:12]


def profiled_traffic(key: str, summary: str = TRAFFIC_SUMMARY, config: str = DEFAULT_TRAFFIC_CONFIG):
    """Bytes per launch measured with the FETCH_SIZE / WRITE_SIZE counters for a workload
    (profiles/), or None when that summary is for another configuration or build."""
    vals = {}
    try:
        for line in open(summary):
            if "=" in line and not line.startswith(("|", "`", " ")):
                k, v = line.strip().split("=", 1)
                vals[k] = v
    except OSError:
        return None
    if vals.get("config") != config or vals.get("sources") != kernel_sources_digest() or key not in vals:
        return None
    return int(float(vals[key]))


def timed_dock(eng, ls, prm, pb, steps, warmup, tag):
    """Warmup + timed docking passes; returns (wall seconds, results, MC launch events, kernel counters)."""
    import torch
    stats = torch.zeros(4, dtype=torch.int64, device="cuda")
    events = []
    for i in range(warmup):
        eng.dock(ls, prm, poses=pb)
        print(f"# {tag} warmup step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    results = None
    t1 = time.perf_counter()
    for i in range(steps):
        results, _ = eng.dock(ls, prm, poses=pb, mc_events=events, mc_stats=stats)
        print(f"# {tag} timed step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    return time.perf_counter() - t1, results, events, stats.cpu().numpy().astype("float64")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(ligands are independent: a multi-GPU run would shard the size-sorted ligand queue "
                          "by estimated cost, one process per GPU, with no collective on the search path)"}))
        sys.exit(2)
    if args.workload == "score":
        return score_microbench(args)
    if args.workload == "copy":
        return copy_probe()
    if args.workload == "gather":
        return gather_microbench(args)
    import numpy as np
    from unidock_engine import _native as N
    from unidock_engine import proposal as LP
    from unidock_engine import synthetic as S
    from unidock_engine.engine import DockEngine, DockParams, heuristic_opt_steps

    N.check(N.lib().ud2_set_device(0), "ud2_set_device")
    t0 = time.time()
    receptor = S.make_receptor()
    box = S.docking_box()
    c4 = args.workload == "c4"
    if c4:
        # 256 chains x 40 steps for both engines; one timed pass each (a pass takes about a minute)
        args.n_lig = 4096 if args.n_lig == 1024 else args.n_lig
        args.nchain = args.learned_nchain = 256
        args.steps, args.warmup, args.no_control = 1, 0, True
        synth = S.config4_ligands(args.n_lig)
    else:
        synth = S.config1_ligands(args.n_lig)
    ligands = [s.ligand for s in synth]
    print(f"# workload generated in {time.time() - t0:.1f}s: {len(ligands)} ligands, "
          f"{sum(l.natom for l in ligands)} atoms, {sum(l.npair for l in ligands)} intra pairs", file=sys.stderr,
          flush=True)
    mode = N.MODE_GRID if args.mode == "grid" else N.MODE_DIRECT
    eng = DockEngine(receptor, box, mode=mode, spacing=S.GRID_SPACING)
    ls = N.LigSet(ligands)
    opt_steps = heuristic_opt_steps(max(l.natom for l in ligands))
    base = dict(refine_steps=5, num_pose=10, rmsd_limit=1.0, energy_range=10.0, seed=12345, randomize=True,
                mode=mode, opt_steps=opt_steps)
    plain_prm = DockParams(nchain=args.nchain, mc_steps=args.mc_steps, search="mc", **base)
    learned = args.search == "learned"
    if learned:
        eng.proposal = N.Mlp(LP.load(args.weights or LP.WEIGHTS))
        lprm = DockParams(nchain=args.learned_nchain, mc_steps=args.learned_steps, search="learned",
                          n_cand=args.n_cand, top_k=args.top_k, sel_temp=args.sel_temp, n_start=args.n_start, **base)
    if c4:
        # warmup on 16 ligands of both atoms-per-lane classes (code objects, first-launch costs)
        by_size = sorted(range(len(ligands)), key=lambda i: ligands[i].natom)
        wls = N.LigSet([ligands[i] for i in by_size[:8] + by_size[-8:]])
        for prm in ([plain_prm, lprm] if learned else [plain_prm]):
            eng.dock(wls, DockParams(**{**prm.__dict__, "mc_steps": 2}))
        print("# c4 warmup done", file=sys.stderr, flush=True)
    # plain engine: one timed pass when it is only the comparison, all passes otherwise
    pb_plain = N.PoseBuffers(ls, args.nchain)
    p_steps, p_warm = (1, 0 if c4 else 1) if learned else (args.steps, args.warmup)
    p_wall, p_res, p_ev, p_st = timed_dock(eng, ls, plain_prm, pb_plain, p_steps, p_warm, "plain")
    p_value = len(ligands) * p_steps / p_wall
    p_top = np.asarray([float(r.energies[0][0]) if r.energies else np.nan for r in p_res])
    control = None
    if learned and not args.no_control:
        # the plain engine at the learned engine's chain / step budget: separates what the
        # network contributes from what the smaller budget does
        cprm = DockParams(nchain=args.learned_nchain, mc_steps=args.learned_steps, search="mc", **base)
        pbc = N.PoseBuffers(ls, args.learned_nchain)
        c_wall, c_res, _, _ = timed_dock(eng, ls, cprm, pbc, 1, 1, "plain-same-budget")
        c_top = np.asarray([float(r.energies[0][0]) if r.energies else np.nan for r in c_res])
        control = {"value": round(len(ligands) / c_wall, 4), "unit": "ligands/s", "nchain": args.learned_nchain,
                   "mc_steps": args.learned_steps, "median_top_affinity": round(float(np.nanmedian(c_top)), 4)}
    if learned:
        pb = N.PoseBuffers(ls, args.learned_nchain)
        wall, res, ev, st = timed_dock(eng, ls, lprm, pb, args.steps, args.warmup, "learned")
        top = np.asarray([float(r.energies[0][0]) if r.energies else np.nan for r in res])
        nchain, msteps, kernel = args.learned_nchain, args.learned_steps, "lmc_kernel<1>"
    else:
        wall, res, ev, st, top = p_wall, p_res, p_ev, p_st, p_top
        nchain, msteps, kernel = args.nchain, args.mc_steps, "mc_kernel<1>"
    if c4:
        kernel = kernel.replace("<1>", "<1> + <2>")  # one search call launches both atoms-per-lane classes
    steps = args.steps
    value = len(ligands) * steps / wall
    mc_ms = np.asarray([a.elapsed_time(b) for a, b in ev])
    launches = len(ev)
    # one search call = one kernel launch per atoms-per-lane class; this batch has one class
    bytes_per_launch = st[1] / max(1, launches)
    achieved = bytes_per_launch / (float(mc_ms.mean()) / 1000.0) / 1e9 if launches else 0.0
    ok = np.isfinite(top) & np.isfinite(p_top)
    shape = "10-80 heavy atoms log-uniform, 0-20 rotatable bonds" if c4 else "10-60 heavy atoms, 0-15 rotatable bonds"
    cfg = {
        "workload": f"{len(ligands)}-ligand uneven batch ({shape}), MC search + "
                    f"cluster + refine + score, {args.mode} inter term",
        "method": (f"learned proposal: {nchain} chains x {msteps} MC steps, each chain started from the best of "
                   f"{args.n_start} randomised starts by the network; per step {args.n_cand} candidate mutations "
                   f"scored by the network, " + (f"one sampled with p ~ exp(-score / {args.sel_temp} kcal/mol)"
                                                 if args.sel_temp > 0 else f"top {args.top_k} rescored exactly") +
                   f", BFGS {opt_steps}"
                   if learned else f"plain MC: {nchain} chains x {msteps} MC steps x BFGS {opt_steps}"),
        "n_ligands": len(ligands),
        "nchain": nchain,
        "mc_steps": msteps,
        "opt_steps": opt_steps,
        "inter_mode": args.mode,
        "mc_ms_per_launch": round(float(mc_ms.mean()), 3) if launches else None,
        "energy_evals_per_step": int(st[0] / steps),
        "evals_per_sec": round(st[0] / wall, 1),
        "mc_steps_per_sec": round(len(ligands) * nchain * msteps * steps / wall, 1),
        "median_top_affinity": round(float(np.nanmedian(top)), 4),
    }
    if learned:
        cfg["plain"] = {
            "value": round(p_value, 4), "unit": "ligands/s", "nchain": args.nchain, "mc_steps": args.mc_steps,
            "ms_per_step": round(p_wall * 1000.0 / p_steps, 3),
            "mc_steps_per_sec": round(len(ligands) * args.nchain * args.mc_steps * p_steps / p_wall, 1),
            "median_top_affinity": round(float(np.nanmedian(p_top)), 4),
        }
        cfg["speedup_vs_plain"] = round(value / p_value, 3)
        cfg["median_top_affinity_diff"] = round(float(np.nanmedian(top) - np.nanmedian(p_top)), 4)
        cfg["median_per_ligand_top_diff"] = round(float(np.median(top[ok] - p_top[ok])), 4)
        cfg["frac_ligands_within_0.1"] = round(float(np.mean(top[ok] - p_top[ok] <= 0.1)), 4)
        if control is not None:
            cfg["plain_same_budget"] = control
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "ligands/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1000.0 / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if not learned else "fp32 (network: bf16 x bf16 -> fp32 MFMA)",
        "data": "synthetic (seeded receptor shell and random torsion-tree ligands)",
        "config": cfg,
        "roofline": {
            "kernel": kernel,
            "bound": "hbm",
            "achieved": round(achieved, 3),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": None,
            "algorithmic_bytes_per_launch": int(bytes_per_launch),
        },
    }
    if args.mode == "grid":
        # the measured ceiling of random 32-byte gathers over the same maps (outside the timed region)
        gc = gather_probe(eng.scene)
        out["roofline"]["gather_ceiling"] = gc
        out["roofline"]["frac_of_gather_ceiling"] = round(achieved / gc["achieved"], 6)
    default_run = (args.n_lig, args.mode, args.nchain, args.mc_steps, args.learned_nchain, args.learned_steps,
                   args.n_cand, args.sel_temp, args.n_start) == (1024, "grid", 128, 40, 36, 40, 16, 1.0, 4)
    if default_run:
        t = profiled_traffic("lmc_traffic_bytes_per_launch") if learned else None
        if t:
            out["roofline"]["traffic"] = t
            out["roofline"]["traffic_source"] = "profiles/lmc_mc_kernels_c1_batch_traffic.md (FETCH_SIZE + WRITE_SIZE)"
    if not args.no_cpu and args.cpu_sample > 0:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(ligands, receptor, box, args.nchain, args.mc_steps, opt_steps,
                                           args.cpu_sample, threads)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
