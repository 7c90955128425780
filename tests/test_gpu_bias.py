"""Bias terms, constraint docking and energy decomposition on the HIP engine vs the
CPU oracle (FP64), through the C ABI, and the reference's bias feature test
through the drop-in pipeline.

Tolerances as for the unbiased path: energies and gradients within 1e-4
relative (floor 1.0); per-atom decomposition terms within 1e-4 absolute plus
1e-4 relative; constraint docking keeps the root fragment bit-exact.
"""
import ctypes as C
import dataclasses
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

VP = C.c_void_p
HERE = os.path.dirname(os.path.abspath(__file__))
PROBES = [2, 3, 4, 5, 6, 10, 11, 15]


def dptr(t):
    return VP(t.data_ptr())


def rel_err(a, b, floor=1.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


@pytest.fixture(scope="module")
def scene(gpu, receptor, box):
    from unidock_engine import synthetic as S
    sc = gpu.Scene(receptor, box)
    sc.build_grids(S.GRID_SPACING, PROBES)
    return sc


@pytest.fixture(scope="module")
def oracle_scenes(scene, receptor, box):
    n, org, sp, nt = scene.grid_info()
    slot = [-1] * 21
    for k, t in enumerate(PROBES):
        slot[t] = k
    return {
        "direct": O.Scene(receptor, box, "direct"),
        "grid": O.Scene(receptor, box, "grid", maps=scene.download_maps(), grid_n=n, grid_origin=org,
                        grid_spacing=sp, slot_of_type=slot),
    }


@pytest.fixture(scope="module")
def lig_b(lig_c2):
    """The c2 ligand with an attractive site near each of three heavy atoms; the first also
    has a second, repulsive site."""
    rng = np.random.default_rng(5)
    heavy = [a for a in range(lig_c2.natom) if lig_c2.types[a] != 0]
    start = np.zeros(lig_c2.natom + 1, np.int32)
    pick = set(int(a) for a in rng.choice(heavy, 3, replace=False))
    rows = []
    for a in range(lig_c2.natom):
        if a in pick:
            rows.append([*(lig_c2.coords[a] + rng.uniform(-1.5, 1.5, 3)), -5.0, 1.5])
            if len(rows) == 1:
                rows.append([*(lig_c2.coords[a] + rng.uniform(-1.5, 1.5, 3)), 2.0, 3.0])
        start[a + 1] = len(rows)
    return dataclasses.replace(lig_c2, bias_start=start, bias=np.asarray(rows, np.float32))


def set_poses(pb, lig, topo, box, nchain, seed):
    from unidock_engine import synthetic as S
    rng = np.random.default_rng(seed)
    pv = S.random_pose_vectors(nchain, lig.ntors, box, seed=seed)
    pv[:, 0:3] = np.asarray(S.POCKET_CENTER, np.float32) + rng.uniform(-4, 4, (nchain, 3))
    starts = []
    for c in range(nchain):
        p = O.pose_from_vec(topo, pv[c].astype(np.float64))
        pb.set_pose(0, c, p.x, p.c, p.rv, p.dihe[:lig.ntors])
        starts.append(O.PoseArr(pb.pose(0, c)["x"], p.c, p.rv, p.dihe[:lig.ntors]))
    return starts


@pytest.mark.parametrize("mode", ["direct", "grid"])
@pytest.mark.parametrize("bias", ["pos", "align"])
def test_energy_grad_with_bias_matches_oracle(gpu, scene, oracle_scenes, lig_b, box, bias, mode):
    import torch
    print(f"energy+gradient, bias {bias}, mode {mode}", flush=True)
    ls = gpu.LigSet([lig_b], bias=bias, bias_k=0.8)
    topo = O.Topo(lig_b, bias=bias, bias_k=0.8)
    nchain = 16
    pb = gpu.PoseBuffers(ls, nchain)
    starts = set_poses(pb, lig_b, topo, box, nchain, 31)
    sel = torch.arange(nchain, dtype=torch.int32, device="cuda")
    m = gpu.MODE_GRID if mode == "grid" else gpu.MODE_DIRECT
    e = torch.zeros(nchain, device="cuda")
    g = torch.zeros((nchain, 64), device="cuda")
    gpu.check(gpu.lib().ud2_energy_grad(scene.h, ls.h, pb.ref, dptr(sel), nchain, m, 1000.0, 1e6, dptr(e), dptr(g),
                                        gpu.stream_ptr()))
    torch.cuda.synchronize()
    dim = 6 + lig_b.ntors
    nbias = 0
    for c in range(nchain):
        e_ref, g_ref, terms = O.energy_grad(topo, starts[c], oracle_scenes[mode], v_cap=1000.0, slope=1e6)
        nbias += abs(terms[3]) > 1e-3
        assert rel_err(float(e[c]), e_ref) < 1e-4, (c, float(e[c]), e_ref)
        scale = max(np.abs(g_ref).max(), 1.0)
        assert np.abs(g[c, :dim].cpu().numpy() - g_ref).max() / scale < 1e-4, c
    assert nbias >= nchain // 2  # the sites are felt


def test_final_score_and_decomposition_match_oracle(gpu, scene, oracle_scenes, lig_b, box):
    import torch
    ls = gpu.LigSet([lig_b], bias="pos", bias_k=1.0)
    topo = O.Topo(lig_b, bias="pos", bias_k=1.0)
    nchain = 8
    pb = gpu.PoseBuffers(ls, nchain)
    set_poses(pb, lig_b, topo, box, nchain, 4)
    sel = torch.arange(nchain, dtype=torch.int32, device="cuda")
    out = torch.zeros((nchain, 4), device="cuda")
    stride = lig_b.natom + 3
    dec = torch.full((nchain, stride, 5), 7.0, device="cuda")
    lib = gpu.lib()
    gpu.check(lib.ud2_final_score(scene.h, ls.h, pb.ref, dptr(sel), nchain, dptr(out), gpu.stream_ptr()))
    gpu.check(lib.ud2_score_decomp(scene.h, ls.h, pb.ref, dptr(sel), nchain, stride, dptr(dec), gpu.stream_ptr()))
    torch.cuda.synchronize()
    dec = dec.cpu().numpy()
    for c in range(nchain):
        x = pb.pose(0, c)["x"]
        ref = O.final_score(topo, x, oracle_scenes["direct"])
        assert rel_err(out[c].cpu().numpy(), ref).max() < 1e-4, (c, out[c], ref)
        dref = O.score_decomp(topo, x, oracle_scenes["direct"])
        d = dec[c, :lig_b.natom]
        assert np.all(np.abs(d - dref) <= 1e-4 + 1e-4 * np.abs(dref)), c
        assert abs(d.sum() - ref[1]) < 1e-3 * max(1.0, abs(ref[1]))
        assert np.all(dec[c, lig_b.natom:] == 7.0)  # rows past natom untouched
    # atom_stride below natom is an argument error, not an out-of-bounds write
    rc = lib.ud2_score_decomp(scene.h, ls.h, pb.ref, dptr(sel), nchain, lig_b.natom - 1, dptr(torch.zeros(1)),
                              gpu.stream_ptr())
    assert rc == 1 and b"atom_stride" in lib.ud2_last_error()


def root_atoms(lig):
    moved = set(int(a) for a in lig.rot_atoms)
    return [a for a in range(lig.natom) if a not in moved]


@pytest.mark.parametrize("search", ["mc", "plain", "exact"])
def test_constraint_search_keeps_the_root_fragment(gpu, scene, lig_c2, search):
    """Constraint docking on the device: randomised, mutated and BFGS-refined poses keep
    the root fragment bit-exact (no rigid step anywhere, bfgs.cuh:571; mc.cu:27-35,159-166),
    and the rigid gradient components are zero (bfgs.cuh:176)."""
    import torch
    print(f"constraint docking, search {search}", flush=True)
    ls = gpu.LigSet([lig_c2], constraint=True)
    nchain = 32
    pb = gpu.PoseBuffers(ls, nchain)
    pb.init()
    mc = gpu.McParams(mc_steps=6, opt_steps=6, randomize=1, seed=77, mode=gpu.MODE_GRID, lig_index_base=0,
                      stats_d=None)
    lib = gpu.lib()
    if search == "mc":
        gpu.check(lib.ud2_mc(scene.h, ls.h, pb.ref, C.byref(mc), gpu.stream_ptr()), "ud2_mc")
    else:
        lp = gpu.LmcParams()
        lp.mc = mc
        lp.policy = gpu.POLICY_PLAIN if search == "plain" else gpu.POLICY_EXACT
        lp.n_cand, lp.top_k, lp.sel_temp, lp.n_start = 8, 1, 0.0, 1
        gpu.check(lib.ud2_lmc(scene.h, ls.h, pb.ref, C.byref(lp), None, gpu.stream_ptr()), "ud2_lmc")
    torch.cuda.synchronize()
    root = root_atoms(lig_c2)
    assert root
    x0 = lig_c2.coords
    e = pb.e[:nchain].cpu().numpy()
    assert np.all(np.isfinite(e)) and np.all(e < 999.0)
    moved = 0
    for c in range(nchain):
        x = pb.pose(0, c)["x"].reshape(-1, 3).astype(np.float32)
        assert np.array_equal(x[root], x0[root]), c
        moved += not np.array_equal(x, x0)
        for a, b in lig_c2.axis:
            d0 = np.linalg.norm(x0[a] - x0[b])
            assert abs(np.linalg.norm(x[a] - x[b]) - d0) < 3e-3 * d0
    assert moved > nchain // 2
    sel = torch.arange(nchain, dtype=torch.int32, device="cuda")
    g = torch.ones((nchain, 64), device="cuda")
    en = torch.zeros(nchain, device="cuda")
    gpu.check(lib.ud2_energy_grad(scene.h, ls.h, pb.ref, dptr(sel), nchain, gpu.MODE_GRID, 1000.0, 1e6, dptr(en),
                                  dptr(g), gpu.stream_ptr()))
    torch.cuda.synchronize()
    g = g.cpu().numpy()
    assert np.all(g[:, :6] == 0) and np.abs(g[:, 6:6 + lig_c2.ntors]).max() > 0


@pytest.mark.parametrize("option", ["constraint", "pos"])
def test_mc_distribution_with_option_matches_oracle(gpu, scene, oracle_scenes, lig_c0, option):
    """MC (randomise + 8 steps x BFGS(8)) in grid mode, 24 seeds x 16 chains on the same Philox
    streams as the FP64 oracle; single trajectories split where float and double round a
    decision differently, so the outcomes are compared as distributions.
    constraint: only the 5 torsions move and nearly every seed finds the same optimum; the
    sorted per-seed best energies agree within 1e-3 kcal/mol at the median and within
    0.05 kcal/mol for every seed (a trajectory that rounds one decision differently can settle
    in a neighbouring minimum; the oracle's own FP32 and FP64 runs agree within 3e-5 on these
    seeds). pos: two atoms pulled to sites 1.5 A off their input positions; the
    best-of-seed is bimodal, so all chains' final energies are pooled: two-sample KS p > 0.01
    and medians within 0.25 kcal/mol."""
    import torch
    from scipy.stats import ks_2samp
    lig = lig_c0
    kw = dict(constraint=True) if option == "constraint" else dict(bias="pos", bias_k=1.0)
    if option == "pos":
        start = np.zeros(lig.natom + 1, np.int32)
        rows = [[*(lig.coords[a] + [1.5, 0.0, 0.0]), -4.0, 1.0] for a in (0, lig.natom - 1)]
        start[1:] = 1
        start[lig.natom] = 2
        lig = dataclasses.replace(lig, bias_start=start, bias=np.asarray(rows, np.float32))
    print(f"MC distribution, {option}", flush=True)
    ls = gpu.LigSet([lig], **kw)
    topo = O.Topo(lig, **kw)
    nchain, seeds = 16, 24
    pb = gpu.PoseBuffers(ls, nchain)
    start_pose = O.PoseArr.initial(lig)
    e_gpu, e_ref = [], []
    for seed in range(seeds):
        pb.init()
        prm = gpu.McParams(mc_steps=8, opt_steps=8, randomize=1, seed=seed, mode=gpu.MODE_GRID, lig_index_base=0)
        gpu.check(gpu.lib().ud2_mc(scene.h, ls.h, pb.ref, C.byref(prm), gpu.stream_ptr()))
        torch.cuda.synchronize()
        e = pb.e[:nchain].cpu().numpy().astype(np.float64)
        assert np.isfinite(e).all()
        res, _, _ = O.mc(topo, oracle_scenes["grid"], start_pose, 0, nchain, 8, 8, seed, True, nthreads=8)
        e_gpu.append(e)
        e_ref.append(res["e"])
    e_gpu, e_ref = np.asarray(e_gpu), np.asarray(e_ref)
    if option == "constraint":
        d = np.abs(np.sort(e_gpu.min(axis=1)) - np.sort(e_ref.min(axis=1)))
        assert np.median(d) < 1e-3 and d.max() < 0.05, d
    else:
        assert ks_2samp(e_gpu.ravel(), e_ref.ravel()).pvalue > 0.01
        assert abs(np.median(e_gpu) - np.median(e_ref)) < 0.25, (np.median(e_gpu), np.median(e_ref))


def test_bias_hbond_feature(gpu, tmp_path):
    """The reference's feature test E/test/feature/bias_hbond (ud2.yaml, validate.py): bias "pos"
    with k = 1, seed 1234567, exhaustiveness 512, 40 MC steps, no refinement, box centre (2, 60, 10)
    size 30; atom 6 of each of the top 5 poses within 0.2 A per coordinate of its site
    (1.782, 66.634, 6.337). Run through the drop-in pipeline."""
    from unidock_engine.api.python import pipeline
    doc = json.load(open(os.path.join(HERE, "golden", "bias_hbond_input.json")))
    pipe = pipeline.DockingPipeline(output_dir=str(tmp_path), center_x=2.0, center_y=60.0, center_z=10.0,
                                    size_x=30.0, size_y=30.0, size_z=30.0, task="screen", search_mode="free",
                                    exhaustiveness=512, randomize=True, mc_steps=40, opt_steps=-1, refine_steps=0,
                                    num_pose=10, rmsd_limit=1.0, energy_range=10.0, seed=1234567, bias="pos",
                                    bias_k=1.0, name_json="input")
    pipe.set_receptor(doc["receptor"])
    pipe.add_ligands({k: v for k, v in doc.items() if k not in ("receptor", "score")})
    pipe.run()
    out = json.load(open(os.path.join(str(tmp_path), "input_1.json")))
    (name, poses), = out.items()
    site = np.asarray([1.782, 66.634, 6.337])
    assert len(poses) >= 5
    for p in poses[:5]:
        xyz = np.asarray(p["coords"][18:21])
        assert np.all(np.abs(xyz - site) <= 0.2), xyz
        assert p["energy"][6] < -10.0  # the well is -15 kcal/mol deep at the site


def test_energy_decomp_through_the_pipeline(gpu, tmp_path):
    """energy_decomp = True (DockTask::run_score, DockTask.cpp:233-266; json.cpp:134-147): every pose
    carries one [gauss1, gauss2, repulsion, hydrophobic, hbond] row per atom; the rows match the
    oracle's score_decomp of the written coordinates and add up to the inter term."""
    from unidock_engine.api.python import pipeline
    from unidock_engine import topology as T
    doc = json.load(open(os.path.join(HERE, "golden", "bias_hbond_input.json")))
    pipe = pipeline.DockingPipeline(output_dir=str(tmp_path), center_x=2.0, center_y=60.0, center_z=10.0,
                                    size_x=30.0, size_y=30.0, size_z=30.0, task="screen", search_mode="free",
                                    exhaustiveness=64, mc_steps=10, refine_steps=1, num_pose=4, seed=3,
                                    energy_decomp=True, name_json="dec")
    pipe.set_receptor(doc["receptor"])
    pipe.add_ligands({"MOL_0": doc["MOL_0"]})
    pipe.run()
    out = json.load(open(os.path.join(str(tmp_path), "dec_1.json")))
    poses = out["MOL_0"]
    box = T.Box.from_center_size((2.0, 60.0, 10.0), (30.0, 30.0, 30.0))
    rec = T.receptor_from_json(doc["receptor"], box)
    lig = T.ligand_from_json("MOL_0", doc["MOL_0"])
    topo, sc = O.Topo(lig), O.Scene(rec, box, "direct")
    for p in poses:
        d = np.asarray(p["decomp"], np.float64)
        assert d.shape == (lig.natom, 5)
        dref = O.score_decomp(topo, np.asarray(p["coords"], np.float64), sc)
        # the JSON keeps 3 decimals of every coordinate and term (truncated): 1e-2 absolute
        assert np.all(np.abs(d - dref) <= 1e-2 + 1e-3 * np.abs(dref))
        assert abs(d.sum() - p["energy"][3]) < 0.25


@pytest.mark.parametrize("option", ["constraint", "pos"])
def test_learned_search_with_option(gpu, scene, receptor, box, lig_b, option):
    """The learned kernel (packaged network, 4 randomised starts pre-screened by the network, 16
    candidates sampled at T = 1) under constraint docking or with bias sites: constraint keeps the
    root fragment bit-exact through the start pre-screen, mutations and BFGS; with bias, every stored
    energy is the exact biased search energy of its pose (V_CAP 1000), re-evaluated separately."""
    import torch
    from unidock_engine import proposal as LP
    print(f"learned search, {option}", flush=True)
    kw = dict(constraint=True) if option == "constraint" else dict(bias="pos", bias_k=1.0)
    ls = gpu.LigSet([lig_b], **kw)
    nchain = 64
    pb = gpu.PoseBuffers(ls, nchain)
    pb.init()
    mlp = gpu.Mlp(LP.load())
    lp = gpu.LmcParams()
    lp.mc = gpu.McParams(mc_steps=4, opt_steps=8, randomize=1, seed=51, mode=gpu.MODE_GRID, lig_index_base=0,
                         stats_d=None)
    lp.policy, lp.n_cand, lp.top_k, lp.sel_temp, lp.n_start = gpu.POLICY_LEARNED, 16, 1, 1.0, 4
    gpu.check(gpu.lib().ud2_lmc(scene.h, ls.h, pb.ref, C.byref(lp), mlp.h, gpu.stream_ptr()), "ud2_lmc")
    torch.cuda.synchronize()
    e = pb.e[:nchain].cpu().numpy()
    assert np.isfinite(e).all()
    if option == "constraint":
        root = root_atoms(lig_b)
        for c in range(nchain):
            x = pb.pose(0, c)["x"].reshape(-1, 3).astype(np.float32)
            assert np.array_equal(x[root], lig_b.coords[root]), c
    else:
        sel = torch.arange(nchain, dtype=torch.int32, device="cuda")
        e2 = torch.zeros(nchain, device="cuda")
        g = torch.zeros((nchain, 64), device="cuda")
        gpu.check(gpu.lib().ud2_energy_grad(scene.h, ls.h, pb.ref, dptr(sel), nchain, gpu.MODE_GRID, 1000.0, 1e6,
                                            dptr(e2), dptr(g), gpu.stream_ptr()))
        torch.cuda.synchronize()
        e2 = e2.cpu().numpy()
        assert np.all(np.abs(e2 - e) <= 1e-4 * np.maximum(1.0, np.abs(e))), np.abs(e2 - e).max()
        topo = O.Topo(lig_b, bias="pos", bias_k=1.0)
        osc = O.Scene(receptor, box, "direct")
        fs = [O.final_score(topo, pb.pose(0, c)["x"], osc)[3] for c in range(nchain)]
        assert min(fs) < -0.1  # the wells are felt by the kept poses
