"""Oracle scoring, gradients, kinematics, line search, BFGS, clustering and
final score pinned by the reference's KATs (E/test/unit/cuda/test_bfgs.cpp),
closed forms, finite differences and invariants. CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O

R_CH = float(np.float32(1.9))


def one_carbon_scene(rec_xyz=((0.0, 0.0, 0.0),), types=(2,)):
    from unidock_engine import topology as T
    rec = T.Receptor(xyz=np.asarray(rec_xyz, np.float32), types=np.asarray(types, np.int32))
    box = T.Box(lo=np.full(3, -30, np.float32), hi=np.full(3, 30, np.float32))
    return rec, box


def atoms_ligand(coords, types, pairs=None):
    from unidock_engine import topology as T
    n = len(coords)
    if pairs is None:
        pairs = [(i, j) for i in range(n) for j in range(i + 1, n)]
    return T.make_ligand("t", np.asarray(coords, np.float32), types, [], [], [], [],
                         np.asarray(pairs, np.int32).reshape(-1, 2))


@pytest.mark.parametrize("r", [0.0, 0.5, 1.0, 1.5])
def test_one_atom_energy_gradient_kat(r):
    """test_bfgs.cpp:84-117: one C_H receptor atom, one C_H ligand atom at (r, 0, 0)."""
    rec, box = one_carbon_scene()
    lig = atoms_ligand([[r, 0, 0]], [2])
    topo, sc = O.Topo(lig), O.Scene(rec, box)
    p = O.PoseArr([r, 0, 0], [r, 0, 0], [0, 0, 0], [])
    e, g, _ = O.energy_grad(topo, p, sc, v_cap=1e30)
    e_ref, g_ref = O.eval_ef(r - 2 * R_CH, 2, 2)
    assert abs(e - e_ref) < 1e-4
    if r == 0.0:
        g_ref = 0.0  # the distance is clamped to epsilon
    assert abs(g[0] - g_ref) < 1e-4 and np.abs(g[1:6]).max() < 1e-4


def test_one_atom_line_search_kat():
    """test_bfgs.cpp:119-130: from r = 1.5 the line search takes alpha = 1 and x_new = x - g."""
    rec, box = one_carbon_scene()
    lig = atoms_ligand([[1.5, 0, 0]], [2])
    topo, sc = O.Topo(lig), O.Scene(rec, box)
    p = O.PoseArr([1.5, 0, 0], [1.5, 0, 0], [0, 0, 0], [])
    e0, g, _ = O.energy_grad(topo, p, sc, v_cap=1e30)
    xn, gn, e1, alpha = O.line_search(topo, sc, p, g, -g, e0, v_cap=1e30)
    assert abs(alpha - 1.0) < 1e-12
    assert abs(xn.x[0] - (1.5 - g[0])) < 1e-4
    r = abs(xn.x[0])
    assert abs(e1 - O.eval_ef(r - 2 * R_CH, 2, 2)[0]) < 1e-4


def test_two_atom_energy_and_fd_gradients_kat():
    """test_bfgs.cpp:214-297: two C_H ligand atoms, one C_H receptor atom; energy = 2 inter + 1
    intra term; translation and rotation gradients against finite differences."""
    rec, box = one_carbon_scene()
    r = 1.0
    ca, cb = np.array([r, 0, 0.0]), np.array([r, 1, 0.0])
    lig = atoms_ligand([ca, cb], [2, 2])
    topo, sc = O.Topo(lig), O.Scene(rec, box)
    c = (ca + cb) / 2
    p = O.PoseArr(np.r_[ca, cb], c, [0, 0, 0], [])
    e, g, _ = O.energy_grad(topo, p, sc, v_cap=1e30)
    e_ref = (O.eval_ef(r - 3.8, 2, 2)[0] + O.eval_ef(math.sqrt(r * r + 1) - 3.8, 2, 2)[0] +
             O.eval_ef(1 - 3.8, 2, 2)[0])
    assert abs(e - e_ref) < 1e-4
    for i in range(3):  # forward differences dr = 1e-3, as the KAT (tolerance 0.01, tightened)
        d = np.zeros(3)
        d[i] = 1e-6
        p2 = O.PoseArr(np.r_[ca + d, cb + d], c + d, [0, 0, 0], [])
        e2, _, _ = O.energy_grad(topo, p2, sc, v_cap=1e30)
        assert abs((e2 - e) / 1e-6 - g[i]) < 1e-4
    rv = np.array([-1.1, 0.2, 0.3])  # rotation gradient at rot_vec = (-1.1, 0.2, 0.3), test_bfgs.cpp:269-292

    def rotated(v):
        q = O.rotvec_to_quat(v)
        xs = [O.quat_rotate(q, a - c) + c for a in (ca, cb)]
        return O.PoseArr(np.concatenate(xs), c, v, [])

    e1, g1, _ = O.energy_grad(topo, rotated(rv), sc, v_cap=1e30)
    for k in range(3):
        dv = np.zeros(3)
        dv[k] = 1e-6
        ep, _, _ = O.energy_grad(topo, rotated(rv + dv), sc, v_cap=1e30)
        em, _, _ = O.energy_grad(topo, rotated(rv - dv), sc, v_cap=1e30)
        assert abs((ep - em) / 2e-6 - g1[3 + k]) < 1e-5


def test_curl_cap_closed_form():
    """Per-atom curl (bfgs.cuh:276-283): e -> e v/(v+e) and force x (v/(v+e))^2 when e > 0."""
    rec, box = one_carbon_scene()
    lig = atoms_ligand([[1.0, 0, 0]], [2])
    topo, sc = O.Topo(lig), O.Scene(rec, box)
    p = O.PoseArr([1.0, 0, 0], [1.0, 0, 0], [0, 0, 0], [])
    e_inf, g_inf, _ = O.energy_grad(topo, p, sc, v_cap=1e30)
    assert e_inf > 0
    for v in (10.0, 1000.0, 0.5):
        e, g, _ = O.energy_grad(topo, p, sc, v_cap=v)
        s = v / (v + e_inf)
        assert abs(e - e_inf * s) < 1e-12 * max(1, e_inf)
        assert abs(g[0] - g_inf[0] * s * s) < 1e-12 * max(1, abs(g_inf[0]))


def test_cutoff_and_penalty_closed_form():
    from unidock_engine import topology as T
    rec, _ = one_carbon_scene()
    box = T.Box(lo=np.asarray([-5, -5, -5], np.float32), hi=np.asarray([20, 5, 5], np.float32))
    sc = O.Scene(rec, box)
    for r, inside in ((7.999, True), (8.0, False), (8.5, False)):
        lig = atoms_ligand([[r, 0, 0]], [2])
        e, _, t = O.energy_forces(O.Topo(lig), [r, 0, 0], sc, v_cap=1e30)
        assert (t[0] != 0.0) == inside  # r^2 < 64 strictly (vina.h:130)
    # penalty: slope x overshoot per heavy atom, gradient +-slope (bfgs.cuh:196-218); hydrogen exempt
    lig = atoms_ligand([[21.0, 6.5, 0.0], [25.0, 0, 0]], [2, 0], pairs=[])
    e, f, t = O.energy_forces(O.Topo(lig), [21.0, 6.5, 0.0, 25.0, 0, 0], sc, v_cap=1e30, slope=100.0)
    assert abs(t[2] - 100.0 * (1.0 + 1.5)) < 1e-9
    assert np.allclose(f[0], [100.0, 100.0, 0.0]) and np.allclose(f[1], 0.0)


def test_intra_pairs_once_equals_symmetric_half():
    """The intra sum visits each pair once; the reference visits the symmetric list and
    halves the energy (bfgs.cuh:305-353). With the curl off both equal the uncurled sum
    of eval_ef over the pairs, which score() computes (score.cpp:21-45)."""
    rng = np.random.default_rng(1)
    xyz = rng.uniform(-3, 3, (7, 3))
    types = [2, 3, 5, 10, 2, 11, 15]
    lig = atoms_ligand(xyz, types)
    rec, box = one_carbon_scene([[50.0, 50.0, 50.0]])
    e, f, t = O.energy_forces(O.Topo(lig), xyz.reshape(-1), O.Scene(rec, box), v_cap=1e30)
    ref = 0.0
    for i in range(7):
        for j in range(i + 1, 7):
            d = np.linalg.norm(xyz[j] - xyz[i])
            if d * d < 64:
                ref += O.eval_ef(d - (O_r(types[i]) + O_r(types[j])), types[i], types[j])[0]
    assert abs(t[1] - ref) < 1e-12
    assert abs(f.sum(axis=0)).max() < 1e-9  # internal forces sum to zero
    # translation / rotation invariance of the intra term
    q = O.rotvec_to_quat([0.4, -1.0, 0.3])
    xyz2 = np.asarray([O.quat_rotate(q, a) for a in xyz]) + np.array([1.0, -2.0, 0.5])
    e2, _, t2 = O.energy_forces(O.Topo(lig), xyz2.reshape(-1), O.Scene(rec, box), v_cap=1e30)
    assert abs(t2[1] - t[1]) < 1e-10


def O_r(t):
    return float(np.float32([1.1, 1.92, 1.9, 1.9, 1.8, 1.8, 1.8, 1.8, 1.7, 1.7, 1.7, 1.7, 1.5, 2.2, 2.1, 2.0, 1.8,
                             2.0, 2.2, 2.3, 1.2][t]))


@pytest.fixture(scope="module")
def c2(receptor, box, lig_c2):
    return O.Topo(lig_c2), O.Scene(receptor, box), lig_c2


def _poses(lig, box, n, seed, inside=True):
    from unidock_engine import synthetic as S
    pv = S.random_pose_vectors(n, lig.ntors, box, seed=seed).astype(np.float64)
    if inside:
        pv[:, 0:3] = np.asarray(S.POCKET_CENTER) + np.random.default_rng(seed).uniform(-6, 6, (n, 3))
    pv[:, 3:7] /= np.linalg.norm(pv[:, 3:7], axis=1, keepdims=True)  # exactly unit in FP64
    return pv


def test_gradient_matches_central_differences(c2, box):
    """Centre, orientation (at the pose's rot_vec) and torsion gradients of E (curl on)."""
    topo, sc, lig = c2
    h = 1e-6
    for pv in _poses(lig, box, 6, 31):
        p = O.pose_from_vec(topo, pv)
        e, g, _ = O.energy_grad(topo, p, sc, v_cap=10.0)
        scale = max(1.0, np.abs(g).max())
        for k in range(3):
            a, b = p.copy(), p.copy()
            a.x.reshape(-1, 3)[:, k] += h
            b.x.reshape(-1, 3)[:, k] -= h
            fd = (O.energy_grad(topo, a, sc, 10.0)[0] - O.energy_grad(topo, b, sc, 10.0)[0]) / (2 * h)
            assert abs(fd - g[k]) < 1e-5 * scale
        xbar = (p.x.reshape(-1, 3) - p.c) @ _rot_matrix(p.rv)  # R^T (x - c), row-wise
        for k in range(3):
            dv = np.zeros(3)
            dv[k] = h

            def at(v):
                q = O.PoseArr((xbar @ _rot_matrix(v).T + p.c).reshape(-1), p.c, v, p.dihe)
                return O.energy_grad(topo, q, sc, 10.0)[0]
            fd = (at(p.rv + dv) - at(p.rv - dv)) / (2 * h)
            assert abs(fd - g[3 + k]) < 1e-5 * scale
        for t in range(lig.ntors):
            a, b = p.copy(), p.copy()
            O.apply_torsion(topo, a, t, h)
            O.apply_torsion(topo, b, t, -h)
            fd = (O.energy_grad(topo, a, sc, 10.0)[0] - O.energy_grad(topo, b, sc, 10.0)[0]) / (2 * h)
            assert abs(fd - g[6 + t]) < 1e-5 * scale


def _rot_matrix(v):
    q = O.rotvec_to_quat(v)
    return np.stack([O.quat_rotate(q, e) for e in np.eye(3)], axis=1)


def test_kinematics_invariants(c2, box):
    topo, sc, lig = c2
    bonds = [(int(a), int(b)) for a, b in lig.axis]
    x0 = lig.coords.astype(np.float64)
    d0 = np.linalg.norm(x0[:, None] - x0[None], axis=-1)
    for pv in _poses(lig, box, 5, 41, inside=False):
        p = O.pose_from_vec(topo, pv)
        x = p.x.reshape(-1, 3)
        for a, b in bonds:
            assert abs(np.linalg.norm(x[a] - x[b]) - d0[a, b]) < 1e-9
        assert np.allclose(x.mean(axis=0), p.c, atol=1e-9)  # centre = centroid
        # torsion steps commute in exact arithmetic (bfgs.cuh:565-567 applies them in order)
        inc = np.random.default_rng(1).uniform(-1, 1, lig.ntors)
        a, b = p.copy(), p.copy()
        for t in range(lig.ntors):
            O.apply_torsion(topo, a, t, inc[t])
        for t in reversed(range(lig.ntors)):
            O.apply_torsion(topo, b, t, inc[t])
        assert np.abs(a.x - b.x).max() < 1e-9
        assert np.allclose(a.dihe, b.dihe)


def test_pose_update_rigid_part(c2):
    """apply_grad_update_pose_tile: centre moves by alpha*g_c, coordinates rotate about the old
    centre by q(alpha*g_rot) and rot_vec becomes alpha*g_rot (bfgs.cuh:571-617)."""
    topo, sc, lig = c2
    p = O.PoseArr.initial(lig)
    g = np.zeros(6 + lig.ntors)
    g[:6] = [0.5, -1.0, 2.0, 0.3, -0.2, 0.4]
    q = p.copy()
    O.apply_update(topo, q, g, 0.5)
    R = _rot_matrix(0.5 * g[3:6])
    # with torsions present the old centre is the centroid refreshed by the torsion step
    cen = p.x.reshape(-1, 3).mean(axis=0)
    exp = (p.x.reshape(-1, 3) - cen) @ R.T + cen + 0.5 * g[:3]
    assert np.abs(q.x.reshape(-1, 3) - exp).max() < 1e-9
    assert np.allclose(q.rv, 0.5 * g[3:6]) and np.allclose(q.c, cen + 0.5 * g[:3], atol=1e-12)


def test_bfgs_pair_minimum_and_concave_quirk():
    """One atom near one receptor atom. From a start in the convex well of eval_ef BFGS converges
    to the minimum along r. From r = 4.6 the profile is concave: the step-0 diagonal rescale
    alpha*(y.p)/(y.y) (bfgs.cuh:800-809) is negative, H stops being positive definite, the search
    ascends, and the restore of bfgs.cuh:818-822 returns the start pose and energy."""
    rec, box = one_carbon_scene()
    rs = np.linspace(3.0, 6.0, 30001)
    emin = min(O.eval_ef(x - 3.8, 2, 2)[0] for x in rs)
    for start, converges in (([3.95, 0.1, -0.05], True), ([4.6, 0.3, -0.2], False)):
        lig = atoms_ligand([start], [2])
        topo, sc = O.Topo(lig), O.Scene(rec, box)
        p = O.PoseArr(start, start, [0, 0, 0], [])
        e0 = O.energy_grad(topo, p, sc, 1000.0)[0]
        O.bfgs(topo, sc, p, 50, v_cap=1000.0)
        assert float(p.energy[0]) <= e0
        if converges:
            assert abs(float(p.energy[0]) - emin) < 1e-6
            assert abs(O.eval_ef(np.linalg.norm(p.x) - 3.8, 2, 2)[1]) < 1e-3
        else:
            assert float(p.energy[0]) == e0 and np.allclose(p.x, start)


def test_bfgs_never_increases_energy(c2, box):
    topo, sc, lig = c2
    for pv in _poses(lig, box, 4, 51, inside=False):
        p = O.pose_from_vec(topo, pv)
        e0 = O.energy_grad(topo, p, sc, 10.0)[0]
        O.bfgs(topo, sc, p, 10, v_cap=10.0)
        assert float(p.energy[0]) <= e0 + 1e-12


def test_grid_maps_are_the_direct_sum(receptor):
    """A map value at a grid point equals the uncurled direct sum for a probe there; the
    trilinear interpolation reproduces grid-point values and is linear along each axis."""
    from unidock_engine import topology as T
    box = T.Box.from_center_size((23.5, -3.0, 7.25), (3.0, 3.0, 3.0))
    probes = [2, 10]
    n = [7, 7, 7]
    maps = O.grid_build(receptor, n, box.lo, 0.5, probes)
    pts = np.asarray([box.lo + 0.5 * np.asarray([i, j, k]) for i, j, k in ((0, 0, 0), (3, 2, 6), (6, 6, 6))],
                     np.float64)
    for s, t in enumerate(probes):
        ref = O.grid_points(receptor, t, pts)
        for (i, j, k), r in zip(((0, 0, 0), (3, 2, 6), (6, 6, 6)), ref):
            assert abs(maps[s, k, j, i] - r) <= 1e-6 * max(1.0, abs(r))
    slot = [-1] * 21
    slot[2], slot[10] = 0, 1
    sc = O.Scene(receptor, box, "grid", maps=maps, grid_n=n, grid_origin=box.lo, grid_spacing=0.5,
                 slot_of_type=slot)
    v, _ = O.grid_eval(sc, 0, pts[1])
    assert abs(v - float(maps[0, 6, 2, 3])) < 1e-9
    a = pts[1] + np.array([0.1, 0, 0])
    b = pts[1] + np.array([0.3, 0, 0])
    m = pts[1] + np.array([0.2, 0, 0])
    va, ga = O.grid_eval(sc, 0, a)
    vb, _ = O.grid_eval(sc, 0, b)
    vm, _ = O.grid_eval(sc, 0, m)
    assert abs(vm - 0.5 * (va + vb)) < 1e-9           # linear inside a cell along x
    assert abs(ga[0] - (vb - va) / 0.2) < 1e-9          # gradient = slope inside the cell


def test_clustering_brute_force():
    """Greedy leaders (cluster.cu:93-129) against a brute-force restatement in FP64 on poses
    whose pairwise RMSDs are well away from the limit."""
    rng = np.random.default_rng(3)
    n, natom = 40, 12
    base = rng.normal(size=(natom, 3)) * 2
    centers = rng.normal(size=(6, natom, 3)) * 3
    xs = np.asarray([base + centers[i % 6] + rng.normal(scale=0.05, size=(natom, 3)) for i in range(n)],
                    np.float32).reshape(n, -1)
    es = rng.normal(size=n).astype(np.float32)
    types = np.asarray([2] * (natom - 2) + [0, 0], np.int32)
    got = O.cluster(es, xs, types, 1.0)
    heavy = types != 0
    order = sorted(range(n), key=lambda i: (es[i], i))
    leaders = []
    for i in order:
        ok = True
        for l in leaders:
            d = xs[i].reshape(-1, 3)[heavy] - xs[l].reshape(-1, 3)[heavy]
            if math.sqrt(float(np.mean(np.sum(d.astype(np.float64) ** 2, axis=1)))) < 1.0:
                ok = False
                break
        if ok:
            leaders.append(i)
    assert list(got) == leaders
    assert len(leaders) == 6


def test_filter_keeps_range_and_count():
    e = np.asarray([3.0, -1.0, 0.5, 9.5, -0.8, 20.0], np.float32)
    cl = np.asarray([0, 1, 2, 3, 4, 5], np.int32)
    out = np.zeros(6, np.int32)
    import ctypes as C
    n = O.lib().ud2o_filter(6, O._p(cl), O._p(e), C.c_float(10.0), 10, O._p(out))
    assert list(out[:n]) == [1, 4, 2, 0]  # within -1 + 10 = 9, ascending
    n = O.lib().ud2o_filter(6, O._p(cl), O._p(e), C.c_float(10.0), 2, O._p(out))
    assert list(out[:n]) == [1, 4]
