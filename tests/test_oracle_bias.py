"""Oracle bias terms, constraint docking and energy decomposition, pinned by the
reference's Z-align KAT (E/test/unit/score/test_bias.cpp:10-28), closed forms of
the two potentials, central differences through the pose gradient, the
invariants of constraint docking and the sum rule of the decomposition. CPU only."""
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle as O


def with_bias(lig, sites):
    """Copy of lig with bias sites {atom: [[x, y, z, V, r2], ...]}."""
    start = np.zeros(lig.natom + 1, np.int32)
    rows = []
    for a in range(lig.natom):
        rows += [list(map(float, s)) for s in sites.get(a, [])]
        start[a + 1] = len(rows)
    return dataclasses.replace(lig, bias_start=start, bias=np.asarray(rows, np.float32).reshape(-1, 5))


def empty_scene():
    from unidock_engine import topology as T
    rec = T.Receptor(xyz=np.zeros((0, 3), np.float32), types=np.zeros(0, np.int32))
    box = T.Box(lo=np.full(3, -50, np.float32), hi=np.full(3, 50, np.float32))
    return O.Scene(rec, box)


def one_atom(xyz, t):
    from unidock_engine import topology as T
    return T.make_ligand("a", np.asarray([xyz], np.float32), [t], [], [], [], [], np.zeros((0, 2), np.int32))


@pytest.mark.parametrize("t,e_exp,f_exp", [(2, 15.0, 22.5), (0, 0.05, 0.075)])
def test_zalign_kat(t, e_exp, f_exp):
    """test_bias.cpp:10-28: site at the origin, a C_H atom at (1, 0, 0), a_qt = 1:
    E = 15, dE/dx = (22.5, 0, 0). A hydrogen takes k_a = 0.1 instead of 30 (vina.h:313)."""
    lig = with_bias(one_atom([1, 0, 0], t), {0: [[0, 0, 0, 1.0, 1.0]]})
    topo = O.Topo(lig, bias="align", bias_k=1.0)
    e, f, terms = O.energy_forces(topo, lig.coords, empty_scene())
    assert abs(e - e_exp) < 1e-9 and abs(terms[3] - e_exp) < 1e-9
    assert np.abs(f[0] - [f_exp, 0, 0]).max() < 1e-9


def test_pos_bias_closed_form():
    """Gaussian well V k exp(-|r|^2 / r2) (vina.h:277-289): V k at the site, V k / e at |r|^2 = r2,
    and a zero bias_k or a "no" bias switches it off."""
    v, k, r2 = -15.0, 0.5, 2.0
    for d2, e_exp in [(0.0, v * k), (r2, v * k / math.e)]:
        lig = with_bias(one_atom([math.sqrt(d2), 0, 0], 6), {0: [[0, 0, 0, v, r2]]})
        e, _, _ = O.energy_forces(O.Topo(lig, bias="pos", bias_k=k), lig.coords, empty_scene())
        assert abs(e - e_exp) < 1e-6  # site and coordinates are float32
    e0, _, _ = O.energy_forces(O.Topo(lig, bias="pos", bias_k=0.0), lig.coords, empty_scene())
    e1, _, _ = O.energy_forces(O.Topo(lig, bias="no", bias_k=k), lig.coords, empty_scene())
    assert e0 == 0.0 and e1 == 0.0


@pytest.mark.parametrize("bias", ["pos", "align"])
def test_bias_forces_are_the_energy_gradient(bias):
    """dE/dx of both potentials against central differences, two sites on one atom."""
    lig = with_bias(one_atom([0.3, -0.4, 0.8], 6), {0: [[0, 0, 0, -15.0, 1.0], [1.0, 0.5, 0.2, 2.0, 3.0]]})
    topo, sc = O.Topo(lig, bias=bias, bias_k=0.7), empty_scene()
    x = lig.coords.astype(np.float64).reshape(-1)
    _, f, _ = O.energy_forces(topo, x, sc)
    h = 1e-6
    for k in range(3):
        a, b = x.copy(), x.copy()
        a[k] += h
        b[k] -= h
        fd = (O.energy_forces(topo, a, sc)[0] - O.energy_forces(topo, b, sc)[0]) / (2 * h)
        assert abs(fd - f[0, k]) < 1e-6 * max(1.0, abs(fd))


@pytest.fixture(scope="module")
def c2(receptor, box, lig_c2):
    return O.Topo(lig_c2), O.Scene(receptor, box), lig_c2


@pytest.fixture(scope="module")
def c2b(receptor, box, lig_c2):
    rng = np.random.default_rng(5)
    heavy = [a for a in range(lig_c2.natom) if lig_c2.types[a] != 0]
    sites = {}
    for a in rng.choice(heavy, 3, replace=False):
        c = lig_c2.coords[a] + rng.uniform(-1.5, 1.5, 3)
        sites[int(a)] = [[*c, -5.0, 1.5]]
    return with_bias(lig_c2, sites), O.Scene(receptor, box)


@pytest.mark.parametrize("bias", ["pos", "align"])
def test_pose_gradient_with_bias_matches_central_differences(c2b, box, bias):
    """Centre and torsion gradients with bias sites on three atoms (the bias block of
    cal_e_f_tile, bfgs.cuh:355-388, flows through cal_grad_tile like any force)."""
    from unidock_engine import synthetic as S
    lig, sc = c2b
    topo = O.Topo(lig, bias=bias, bias_k=0.8)
    h = 1e-6
    pvs = S.random_pose_vectors(3, lig.ntors, box, seed=17).astype(np.float64)
    pvs[:, 0:3] = np.asarray(S.POCKET_CENTER)
    pvs[:, 3:7] /= np.linalg.norm(pvs[:, 3:7], axis=1, keepdims=True)
    for pv in pvs:
        p = O.pose_from_vec(topo, pv)
        e, g, terms = O.energy_grad(topo, p, sc, v_cap=10.0)
        assert abs(e - terms.sum()) < 1e-9 * max(1.0, abs(e)) and terms[3] != 0.0
        scale = max(1.0, np.abs(g).max())
        for k in range(3):
            a, b = p.copy(), p.copy()
            a.x.reshape(-1, 3)[:, k] += h
            b.x.reshape(-1, 3)[:, k] -= h
            fd = (O.energy_grad(topo, a, sc, 10.0)[0] - O.energy_grad(topo, b, sc, 10.0)[0]) / (2 * h)
            assert abs(fd - g[k]) < 1e-5 * scale
        for t in range(lig.ntors):
            a, b = p.copy(), p.copy()
            O.apply_torsion(topo, a, t, h)
            O.apply_torsion(topo, b, t, -h)
            fd = (O.energy_grad(topo, a, sc, 10.0)[0] - O.energy_grad(topo, b, sc, 10.0)[0]) / (2 * h)
            assert abs(fd - g[6 + t]) < 1e-5 * scale


def test_final_score_reports_bias_outside_the_total(c2b):
    """score() (score.cpp:85-107) puts the bias in its own slot; energy_forces adds it to E."""
    lig, sc = c2b
    topo = O.Topo(lig, bias="pos", bias_k=1.0)
    x = lig.coords.reshape(-1).astype(np.float64)
    fs = O.final_score(topo, x, sc)
    _, _, terms = O.energy_forces(topo, x, sc, v_cap=1e30)
    assert abs(fs[3] - terms[3]) < 1e-12 and fs[3] < 0
    fs0 = O.final_score(O.Topo(lig), x, sc)
    assert np.array_equal(fs[:3], fs0[:3]) and fs0[3] == 0.0


def root_atoms(lig):
    moved = set(int(a) for a in lig.rot_atoms)
    return [a for a in range(lig.natom) if a not in moved]


def test_constraint_gradient_and_update(c2):
    """Constraint docking: no rigid gradient (bfgs.cuh:176), no rigid step (bfgs.cuh:571);
    torsion components unchanged; atoms that no torsion moves stay where they are."""
    topo, sc, lig = c2
    tc = O.Topo(lig, constraint=True)
    p = O.PoseArr.initial(lig)
    e, g, _ = O.energy_grad(topo, p, sc)
    ec, gc, _ = O.energy_grad(tc, p, sc)
    assert e == ec and np.all(gc[:6] == 0) and np.array_equal(gc[6:], g[6:])
    step = np.random.default_rng(3).uniform(-1, 1, 6 + lig.ntors)
    q = p.copy()
    O.apply_update(tc, q, step, 0.7)
    root = root_atoms(lig)
    assert root and np.array_equal(q.x.reshape(-1, 3)[root], p.x.reshape(-1, 3)[root])
    assert np.array_equal(q.rv, p.rv)


def test_constraint_mc_keeps_the_root_fragment(c2):
    """MC + BFGS under constraint docking (mc.cu:159-166; randomisation keeps the rigid pose,
    mc.cu:27-35): the root fragment never moves, bond lengths hold, the search still improves E."""
    topo, sc, lig = c2
    tc = O.Topo(lig, constraint=True)
    start = O.PoseArr.initial(lig)
    out, _, _ = O.mc(tc, sc, start, lig_index=0, nchain=4, mc_steps=6, opt_steps=4, randomize_flag=True, seed=9)
    root = root_atoms(lig)
    x0 = lig.coords.astype(np.float64)
    for c in range(4):
        x = out["x"][c].reshape(-1, 3)
        assert np.array_equal(x[root], x0[root])
        for a, b in lig.axis:
            assert abs(np.linalg.norm(x[a] - x[b]) - np.linalg.norm(x0[a] - x0[b])) < 1e-9
    e0 = O.energy_grad(tc, start, sc)[0]
    assert out["e"].min() < e0


def test_constraint_mutation_of_a_rigid_ligand_is_a_no_op():
    """mc.cu:159-166: with no torsion the constraint mutation index is past every DOF."""
    from unidock_engine import topology as T
    lig = T.make_ligand("r", np.asarray([[0, 0, 0], [1.5, 0, 0], [1.5, 1.5, 0]], np.float32), [2, 2, 6],
                        [], [], [], [], np.zeros((0, 2), np.int32))
    sc = empty_scene()
    for seed in range(5):
        p = O.PoseArr.initial(lig)
        O.mutate(O.Topo(lig, constraint=True), sc, p, seed, 0, 0)
        assert np.array_equal(p.x, O.PoseArr.initial(lig).x)
        q = O.PoseArr.initial(lig)
        O.mutate(O.Topo(lig), sc, q, seed, 0, 0)
    assert not np.array_equal(q.x, O.PoseArr.initial(lig).x)


def test_randomize_keeps_what_the_options_fix():
    """randomize_pose_tile (mc.cu:18-104) on a rigid ligand: constraint docking keeps the pose,
    the align bias keeps the centre but not the orientation, neither keeps both."""
    from unidock_engine import topology as T
    lig = T.make_ligand("r", np.asarray([[0, 0, 0], [1.5, 0, 0], [1.5, 1.5, 0]], np.float32), [2, 2, 6],
                        [], [], [], [], np.zeros((0, 2), np.int32))
    sc = empty_scene()
    runs = {}
    for name, kw in [("free", {}), ("constraint", dict(constraint=True)), ("align", dict(bias="align"))]:
        p = O.PoseArr.initial(lig)
        O.randomize(O.Topo(lig, **kw), sc, p, 12345, 0, 3, 8)
        runs[name] = p
    p0 = O.PoseArr.initial(lig)
    assert np.array_equal(runs["constraint"].x, p0.x)
    assert np.allclose(runs["align"].c, p0.c, atol=1e-12)
    assert np.allclose(runs["align"].x.reshape(-1, 3).mean(axis=0), p0.c, atol=1e-9)
    assert not np.allclose(runs["align"].x, p0.x)
    assert not np.allclose(runs["free"].c, p0.c)


def test_decomposition_sums_to_the_inter_term(c2):
    """score_decomp (score.cpp:115-147): per-atom rows of the five weighted terms add up to
    the inter term of score(); hydrogen rows are zero; every term agrees with eval_ef's split."""
    topo, sc, lig = c2
    from unidock_engine import synthetic as S
    pvs = S.random_pose_vectors(3, lig.ntors, S.docking_box(), seed=23).astype(np.float64)
    pvs[:, 0:3] = np.asarray(S.POCKET_CENTER)
    pvs[:, 3:7] /= np.linalg.norm(pvs[:, 3:7], axis=1, keepdims=True)
    for pv in pvs:
        p = O.pose_from_vec(topo, pv)
        d = O.score_decomp(topo, p.x, sc)
        fs = O.final_score(topo, p.x, sc)
        assert abs(d.sum() - fs[1]) < 1e-9 * max(1.0, abs(fs[1]))
        assert np.all(d[lig.types == 0] == 0)
        assert np.all(d[:, 0] <= 0) and np.all(d[:, 1] <= 0) and np.all(d[:, 2] >= 0)
        assert np.all(d[:, 3] <= 0) and np.all(d[:, 4] <= 0)
