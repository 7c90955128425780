"""Seeded synthetic docking workloads: a receptor shell around a pocket and
random torsion-tree ligands.

Receptor: heavy atoms with XS types drawn from a protein-like element mix
(C 62 %, N 17 %, O 19 %, S 2 %; XS subtypes split as in the 5S8I receptor:
C_H 0.466 / C_P 0.534, N_D 0.952 / N_P 0.036 / N_A 0.012, O_A 0.838 /
O_DA 0.162, S -> S_P), placed by random sequential addition at >= 3.0 A from
each other in a shell (inner radius 16 A, outer 30 A) around the pocket
centre, the core left empty.

Ligand: a random tree of heavy atoms (each new atom bonds to a random earlier
atom of degree < 4) with 1.5 A bonds, 109.5 deg bond angles and random
dihedrals, non-bonded contacts kept >= 2.0 A; the rotatable bonds are drawn
among the non-terminal edges; the root is the largest rigid fragment; each
torsion rotates the atoms on the far side of its bond ("to" atom included);
1-2/1-3 and 1-4 partner lists follow the graph; torsion ranges are the full
circle. The ligand is returned both as an engine JSON object and as the
engine's :class:`Ligand` (built through the same preprocessing as JSON input).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .topology import (Box, Ligand, Receptor, XS_C_H, XS_C_P, XS_N_A, XS_N_D, XS_N_P, XS_O_A, XS_O_DA, XS_S_P,
                       ligand_from_json)

POCKET_CENTER = (12.5, -3.0, 7.25)
BOX_SIZE = 30.0
GRID_SPACING = 0.375

_ELEMENTS = ("C", "N", "O", "S")
_SUBTYPES = {
    "C": ((XS_C_H, XS_C_P), (0.466, 0.534)),
    "N": ((XS_N_D, XS_N_P, XS_N_A), (0.952, 0.036, 0.012)),
    "O": ((XS_O_A, XS_O_DA), (0.838, 0.162)),
    "S": ((XS_S_P,), (1.0,)),
}
# forcefield type ids only have to round-trip through the JSON; one per element
_FF_TYPE = {"C": 3, "N": 50, "O": 4, "S": 44}


def _draw_types(rng: np.random.Generator, n: int, element_p) -> tuple[np.ndarray, list]:
    elems = rng.choice(len(_ELEMENTS), size=n, p=np.asarray(element_p) / np.sum(element_p))
    types = np.empty(n, dtype=np.int32)
    names = []
    for i, e in enumerate(elems):
        el = _ELEMENTS[e]
        sub, p = _SUBTYPES[el]
        types[i] = sub[rng.choice(len(sub), p=np.asarray(p) / np.sum(p))]
        names.append(el)
    return types, names


def make_receptor(n_atoms: int = 2000, seed: int = 1, center=POCKET_CENTER, r_in: float = 16.0,
                  r_out: float = 30.0, min_dist: float = 3.0) -> Receptor:
    """Random-sequential-addition receptor shell (seed 1 for every config)."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, dtype=np.float64)
    cell = min_dist
    grid: dict = {}
    pts = []
    tries = 0
    while len(pts) < n_atoms:
        tries += 1
        if tries > 200 * n_atoms:
            raise RuntimeError("receptor placement did not converge")
        # uniform in the shell volume
        u = rng.random()
        r = (r_in ** 3 + u * (r_out ** 3 - r_in ** 3)) ** (1.0 / 3.0)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        p = c + r * v
        key = tuple(np.floor(p / cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if np.sum((q - p) ** 2) < min_dist * min_dist:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(key, []).append(p)
            pts.append(p)
    xyz = np.asarray(pts, dtype=np.float32)
    types, names = _draw_types(rng, n_atoms, (0.62, 0.17, 0.19, 0.02))
    ff = np.asarray([_FF_TYPE[e] for e in names], dtype=np.int32)
    q = np.zeros(n_atoms, dtype=np.float32)
    return Receptor(xyz=xyz, types=types, ff_types=ff, charges=q)


def docking_box(center=POCKET_CENTER, size: float = BOX_SIZE) -> Box:
    return Box.from_center_size(center, (size, size, size))


def receptor_json(rec: Receptor) -> list:
    return [[float(x), float(y), float(z), int(t), int(f), float(q)]
            for (x, y, z), t, f, q in zip(rec.xyz, rec.types, rec.ff_types, rec.charges)]


# ------------------------------------------------------------------ ligands
def _perp(v: np.ndarray) -> np.ndarray:
    a = np.array([1.0, 0.0, 0.0]) if abs(v[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    p = np.cross(v, a)
    return p / np.linalg.norm(p)


def _place(parent_xyz, grand_xyz, great_xyz, rng, bond=1.5, angle_deg=109.5):
    """New atom bonded to parent at 1.5 A, angle 109.5 deg to grand, random dihedral about grand-parent."""
    theta = math.radians(angle_deg)
    if grand_xyz is None:
        d = rng.normal(size=3)
        return parent_xyz + bond * d / np.linalg.norm(d)
    b = parent_xyz - grand_xyz
    b /= np.linalg.norm(b)
    if great_xyz is not None:
        n = np.cross(grand_xyz - great_xyz, b)
        if np.linalg.norm(n) < 1e-6:
            n = _perp(b)
        n /= np.linalg.norm(n)
    else:
        n = _perp(b)
    m = np.cross(n, b)
    phi = rng.uniform(-math.pi, math.pi)
    d = -b * math.cos(theta) + math.sin(theta) * (math.cos(phi) * m + math.sin(phi) * n)
    return parent_xyz + bond * d


def _dihedral_deg(p0, p1, p2, p3) -> float:
    b0 = p0 - p1
    b1 = p2 - p1
    b2 = p3 - p2
    b1n = b1 / np.linalg.norm(b1)
    v = b0 - np.dot(b0, b1n) * b1n
    w = b2 - np.dot(b2, b1n) * b1n
    x = np.dot(v, w)
    y = np.dot(np.cross(b1n, v), w)
    return math.degrees(math.atan2(y, x))


@dataclass
class SynthLigand:
    json_obj: dict
    ligand: Ligand
    seed: int


def make_ligand(seed: int, n_heavy: int, n_rot: int, center=POCKET_CENTER, name: str | None = None,
                max_tries: int = 200) -> SynthLigand:
    """Random torsion-tree ligand; ``n_rot`` is capped at the number of non-terminal bonds."""
    rng = np.random.default_rng(seed)
    xyz = np.zeros((n_heavy, 3))
    parent = [-1] * n_heavy
    nbrs: list[list[int]] = [[] for _ in range(n_heavy)]
    for i in range(1, n_heavy):
        placed = False
        for _ in range(max_tries):
            cand = [j for j in range(i) if len(nbrs[j]) < 4]
            par = int(rng.choice(cand))
            grand = nbrs[par][0] if nbrs[par] else None
            if grand is None and par > 0:
                grand = parent[par]
            great = None
            if grand is not None:
                gg = [k for k in nbrs[grand] if k != par]
                great = gg[0] if gg else None
            p = _place(xyz[par], None if grand is None else xyz[grand], None if great is None else xyz[great], rng)
            # non-bonded contacts >= 2.0 A (the parent is the only bonded partner)
            d2 = np.sum((xyz[:i] - p) ** 2, axis=1)
            d2[par] = 99.0
            if np.all(d2 >= 4.0):
                xyz[i] = p
                parent[i] = par
                nbrs[par].append(i)
                nbrs[i].append(par)
                placed = True
                break
        if not placed:
            raise RuntimeError(f"ligand seed {seed}: could not place atom {i}")
    types, names = _draw_types(rng, n_heavy, (0.70, 0.14, 0.14, 0.02))
    edges = [(parent[i], i) for i in range(1, n_heavy)]
    nonterminal = [e for e in edges if len(nbrs[e[0]]) > 1 and len(nbrs[e[1]]) > 1]
    n_rot = min(n_rot, len(nonterminal))
    rot_idx = sorted(rng.choice(len(nonterminal), size=n_rot, replace=False).tolist()) if n_rot else []
    rot_edges = [nonterminal[k] for k in rot_idx]
    rot_set = {frozenset(e) for e in rot_edges}

    # rigid fragments after cutting the rotatable bonds
    frag_id = [-1] * n_heavy
    frags = []
    for s in range(n_heavy):
        if frag_id[s] >= 0:
            continue
        stack = [s]
        frag_id[s] = len(frags)
        members = []
        while stack:
            a = stack.pop()
            members.append(a)
            for b in nbrs[a]:
                if frag_id[b] < 0 and frozenset((a, b)) not in rot_set:
                    frag_id[b] = len(frags)
                    stack.append(b)
        frags.append(sorted(members))
    root_frag = max(range(len(frags)), key=lambda k: (len(frags[k]), -min(frags[k])))
    root_atoms = frags[root_frag]

    # orient each rotatable bond away from the root: BFS over the atom graph from the root fragment
    depth = [-1] * n_heavy
    queue = list(root_atoms)
    for a in root_atoms:
        depth[a] = 0
    order = []
    while queue:
        a = queue.pop(0)
        order.append(a)
        for b in nbrs[a]:
            if depth[b] < 0:
                depth[b] = depth[a] + 1
                queue.append(b)
    torsions = []
    for (u, v) in rot_edges:
        frm, to = (u, v) if depth[u] < depth[v] else (v, u)
        # far side of the bond: atoms reachable from "to" without crossing back over (from, to)
        seen = {to}
        stack = [to]
        while stack:
            a = stack.pop()
            for b in nbrs[a]:
                if b == frm and a == to:
                    continue
                if b not in seen:
                    seen.add(b)
                    stack.append(b)
        a0 = [k for k in nbrs[frm] if k != to][0]
        a3 = [k for k in nbrs[to] if k != frm][0]
        torsions.append((depth[to], frm, to, a0, a3, sorted(seen)))
    torsions.sort(key=lambda t: (t[0], t[2]))

    # centre the conformer on the pocket centre
    xyz = xyz - xyz.mean(axis=0) + np.asarray(center)
    xyz32 = xyz.astype(np.float32)

    # graph partner lists, P/ligand_topology/utils.py:411-465 semantics
    def within(a, k):
        dist = {a: 0}
        frontier = [a]
        for d in range(1, k + 1):
            nxt = []
            for x in frontier:
                for y in nbrs[x]:
                    if y not in dist:
                        dist[y] = d
                        nxt.append(y)
            frontier = nxt
        return dist

    atoms = []
    for a in range(n_heavy):
        dist = within(a, 3)
        l1213 = sorted(b for b, d in dist.items() if d in (1, 2))
        l14 = sorted(b for b, d in dist.items() if d == 3)
        atoms.append([float(xyz32[a, 0]), float(xyz32[a, 1]), float(xyz32[a, 2]), int(types[a]),
                      _FF_TYPE[names[a]], 0.0, l1213, l14])
    tors_json = []
    for (_, frm, to, a0, a3, rot) in torsions:
        deg = _dihedral_deg(xyz32[a0].astype(np.float64), xyz32[frm].astype(np.float64),
                            xyz32[to].astype(np.float64), xyz32[a3].astype(np.float64))
        tors_json.append([[a0, frm, to, a3], deg, [[-180.0, 180.0]], rot, [[1, 0.0, 1, 0.0]]])
    obj = {"atoms": atoms, "torsions": tors_json, "root_atoms": root_atoms,
           "fragment_atom_idx": [frags[k] for k in range(len(frags))]}
    nm = name or f"lig{seed}"
    return SynthLigand(json_obj=obj, ligand=ligand_from_json(nm, obj, use_tor_lib=False), seed=seed)


# ------------------------------------------------------------- workloads
def config0_ligand() -> SynthLigand:
    """Small plain-method case: 24 heavy atoms, 5 rotatable bonds, seed 1000."""
    return make_ligand(1000, 24, 5)


def config1_ligands(n: int = 1024, base_seed: int = 2000) -> list[SynthLigand]:
    """Uneven batch: heavy atoms uniform in 10-60, rotatable bonds uniform in 0-15, seeds 2000+i."""
    out = []
    for i in range(n):
        r = np.random.default_rng(base_seed + i + 7_000_000)
        na = int(r.integers(10, 61))
        nr = int(r.integers(0, 16))
        out.append(make_ligand(base_seed + i, na, nr))
    return out


def config2_ligand() -> SynthLigand:
    """Scoring microbenchmark ligand: 40 heavy atoms, 8 rotatable bonds, seed 3000."""
    return make_ligand(3000, 40, 8)


def config4_ligands(n: int = 4096, base_seed: int = 4000) -> list[SynthLigand]:
    """Large uneven batch: heavy atoms 10-80 log-uniform, rotatable bonds 0-20, seeds 4000+i."""
    out = []
    for i in range(n):
        r = np.random.default_rng(base_seed + i + 7_000_000)
        na = int(round(math.exp(r.uniform(math.log(10), math.log(80)))))
        nr = int(r.integers(0, 21))
        out.append(make_ligand(base_seed + i, na, nr))
    return out


def random_pose_vectors(n: int, ntors: int, box: Box, seed: int = 3, canonical_w: bool = True) -> np.ndarray:
    """Pose vectors [centre(3), quaternion (w,x,y,z), torsions]: centre uniform in the box, rotation
    uniform on SO(3) (Shoemake's method), torsions uniform in [-pi, pi). With ``canonical_w`` the
    quaternion is flipped to w >= 0 so that its rotation angle is at most pi."""
    rng = np.random.default_rng(seed)
    pv = np.empty((n, 7 + ntors), dtype=np.float32)
    pv[:, 0:3] = box.lo + (box.hi - box.lo) * rng.random((n, 3))
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3)], axis=1)
    if canonical_w:
        q[q[:, 0] < 0] *= -1.0
    pv[:, 3:7] = q
    pv[:, 7:] = rng.uniform(-np.pi, np.pi, size=(n, ntors))
    return pv
