"""Ligand and receptor data model of the engine, and the topology preprocessing
that turns the engine's JSON input into it.

The JSON input contract and its semantics follow the reference parser
(``unidock/unidock_engine/src/format/rapidjson_parser.cpp:25-296`` and
``src/format/json.cpp:34-76``): hydrogen receptor atoms are dropped and the
receptor is cropped to the docking box grown by the 8 A cutoff; each torsion's
axis is its two middle atoms; torsion ranges are the full circle unless the
torsion library is on, in which case a range that crosses +-180 deg is split in
two; intra-molecular pairs are all atom pairs that are neither 1-2/1-3 nor 1-4
partners and do not share a rigid fragment, with fragments built by
``split_torsions_into_frags`` (``rapidjson_parser.cpp:227-296``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

# XS atom types (src/constants/constants.h:48-68).
XS_H, XS_B, XS_C_H, XS_C_P, XS_N_P, XS_N_D, XS_N_A, XS_N_DA = 0, 1, 2, 3, 4, 5, 6, 7
XS_O_P, XS_O_D, XS_O_A, XS_O_DA, XS_F_H, XS_Si, XS_P_P, XS_S_P = 8, 9, 10, 11, 12, 13, 14, 15
XS_Cl_H, XS_Br_H, XS_I_H, XS_At, XS_Met_D = 16, 17, 18, 19, 20
XS_NTYPES = 21

# float32 value of the reference's PI = Real(3.1415927) (src/myutils/common.h:45).
PI_F32 = float(np.float32(3.1415927))
DEG2RAD_F32 = np.float32(np.float32(PI_F32) / np.float32(180.0))
RAD2DEG_F32 = np.float32(np.float32(180.0) / np.float32(PI_F32))
VINA_CUTOFF = 8.0  # src/constants/constants.h:9


@dataclass
class Box:
    lo: np.ndarray  # (3,) float32
    hi: np.ndarray  # (3,) float32

    @staticmethod
    def from_center_size(center: Sequence[float], size: Sequence[float]) -> "Box":
        """Box of DockingPipeline (api/python/pipeline.cpp:88-93): lo = c - s/2, hi = c + s/2 in float."""
        c = np.asarray(center, dtype=np.float32)
        s = np.asarray(size, dtype=np.float32)
        return Box(lo=(c - s / np.float32(2)).astype(np.float32), hi=(c + s / np.float32(2)).astype(np.float32))

    def grown(self, margin: float) -> "Box":
        m = np.float32(margin)
        return Box(lo=(self.lo - m).astype(np.float32), hi=(self.hi + m).astype(np.float32))

    def contains(self, xyz: np.ndarray) -> np.ndarray:
        return np.all((xyz >= self.lo) & (xyz <= self.hi), axis=1)


@dataclass
class Receptor:
    xyz: np.ndarray    # (n, 3) float32, heavy atoms only
    types: np.ndarray  # (n,) int32 XS types
    ff_types: np.ndarray | None = None
    charges: np.ndarray | None = None

    @property
    def natom(self) -> int:
        return int(self.xyz.shape[0])


@dataclass
class Ligand:
    """Topology and reference conformation of one flexible ligand.

    Field meaning follows FlexTopo / FlexParamVina (src/model/model.h:95-108,
    src/score/vina.h:17-25): ``axis[t] = (from, to)``, ``rot_atoms`` lists the
    atoms each torsion rotates (its "to" atom included), ``ranges`` are the
    allowed dihedral intervals in radians, ``pairs`` the intra-molecular pairs
    (i < j) that are scored.
    """

    name: str
    coords: np.ndarray       # (n, 3) float32
    types: np.ndarray        # (n,) int32
    dihedrals: np.ndarray    # (nt,) float32 radians
    axis: np.ndarray         # (nt, 2) int32
    rot_start: np.ndarray    # (nt+1,) int32
    rot_atoms: np.ndarray    # (R,) int32
    range_start: np.ndarray  # (nt+1,) int32, in units of (lo, hi) pairs
    ranges: np.ndarray       # (Q, 2) float32
    pairs: np.ndarray        # (P, 2) int32
    ff_types: np.ndarray | None = None
    charges: np.ndarray | None = None
    # bias sites: atom a owns rows bias_start[a] .. bias_start[a+1]-1 of bias (x, y, z, V, r2)
    bias_start: np.ndarray | None = None  # (n+1,) int32, or None: no sites
    bias: np.ndarray | None = None        # (B, 5) float32
    json_obj: dict | None = field(default=None, repr=False)
    pairs_1213: set | None = field(default=None, repr=False)
    pairs_14: set | None = field(default=None, repr=False)
    frags: list | None = field(default=None, repr=False)

    @property
    def natom(self) -> int:
        return int(self.coords.shape[0])

    @property
    def ntors(self) -> int:
        return int(self.axis.shape[0])

    @property
    def npair(self) -> int:
        return int(self.pairs.shape[0])

    @property
    def n_heavy(self) -> int:
        return int(np.count_nonzero(self.types != XS_H))

    def centroid(self) -> np.ndarray:
        """Input centre as the reference parser sums it (rapidjson_parser.cpp:67-82,197-199): float sums."""
        s = np.zeros(3, dtype=np.float32)
        for a in range(self.natom):
            s = (s + self.coords[a]).astype(np.float32)
        return (s / np.float32(self.natom)).astype(np.float32)


# ----------------------------------------------------------------- fragments
def split_torsions_into_frags(root: Iterable[int], torsions: Sequence[tuple[tuple[int, int], Sequence[int]]]):
    """Rigid fragments from the root atoms and the torsions.

    Restates ``RapidJsonParser::split_torsions_into_frags``
    (src/format/rapidjson_parser.cpp:227-296): torsions touching the root make
    their own fragment (rotated atoms + axis) and add their axis to the root
    fragment; every other torsion then splits the first fragment that holds
    both of its axis atoms into (fragment & rotated) and (fragment - rotated),
    both with the axis added. ``torsions`` holds ((from, to), rotated_atoms).
    """
    root = set(int(a) for a in root)
    frag0 = set(root)
    frags: list[set[int]] = []
    remaining = list(range(len(torsions)))
    for it in list(remaining):
        (a0, a1), rot = torsions[it]
        if a0 in root or a1 in root:
            remaining.remove(it)
            frag0.add(a0)
            frag0.add(a1)
            f = set(int(a) for a in rot)
            f.add(a0)
            f.add(a1)
            frags.append(f)
    frags.append(frag0)
    for it in remaining:
        (a0, a1), rot = torsions[it]
        for k, frag in enumerate(frags):
            if a0 in frag and a1 in frag:
                rs = set(int(a) for a in rot)
                f1 = frag & rs
                f2 = frag - rs
                f1 |= {a0, a1}
                f2 |= {a0, a1}
                frags[k] = f1
                frags.append(f2)
                break
    return frags


def intra_pairs(natom: int, pairs_1213: set, pairs_14: set, frags: list) -> np.ndarray:
    """Scored intra pairs (i < j), rapidjson_parser.cpp:201-221."""
    frag_of = [[] for _ in range(natom)]
    for k, f in enumerate(frags):
        for a in f:
            if 0 <= a < natom:
                frag_of[a].append(k)
    out = []
    for i in range(natom):
        fi = set(frag_of[i])
        for j in range(i + 1, natom):
            if (i, j) in pairs_1213 or (i, j) in pairs_14:
                continue
            if fi and fi.intersection(frag_of[j]):
                continue
            out.append((i, j))
    return np.asarray(out, dtype=np.int32).reshape(-1, 2)


def torsion_ranges(range_deg: Sequence[Sequence[float]], use_tor_lib: bool) -> list[tuple[float, float]]:
    """Allowed dihedral intervals in radians (rapidjson_parser.cpp:151-173)."""
    if not use_tor_lib:
        return [(-PI_F32, PI_F32)]
    out = []
    for lo_d, hi_d in range_deg:
        lo = float(np.float32(np.float32(lo_d) * DEG2RAD_F32))
        hi = float(np.float32(np.float32(hi_d) * DEG2RAD_F32))
        if lo < hi:
            out.append((lo, hi))
        elif lo > 0 and hi < 0:
            out.append((lo, PI_F32))
            out.append((-PI_F32, hi))
        else:
            raise ValueError(f"wrong torsion range, lo > hi: {lo_d}, {hi_d}")
    return out


def ligand_from_json(name: str, obj: dict, use_tor_lib: bool = False) -> Ligand:
    """One ligand of the engine JSON (rapidjson_parser.cpp:54-225)."""
    atoms = obj["atoms"]
    n = len(atoms)
    coords = np.asarray([[a[0], a[1], a[2]] for a in atoms], dtype=np.float32).reshape(n, 3)
    types = np.asarray([a[3] for a in atoms], dtype=np.int32)
    ff = np.asarray([a[4] for a in atoms], dtype=np.int32)
    q = np.asarray([a[5] for a in atoms], dtype=np.float32)
    p1213: set = set()
    p14: set = set()
    for ia, a in enumerate(atoms):
        for ib in a[6]:
            ib = int(ib)
            if ib != ia:
                p1213.add((min(ia, ib), max(ia, ib)))
    for ia, a in enumerate(atoms):
        for ib in a[7]:
            ib = int(ib)
            if ib != ia:
                p = (min(ia, ib), max(ia, ib))
                if p not in p1213:
                    p14.add(p)
    torsions = obj.get("torsions", [])
    axis, dihe, rot_lists, range_lists = [], [], [], []
    for tor in torsions:
        four = [int(v) for v in tor[0]]
        axis.append((four[1], four[2]))
        dihe.append(np.float32(np.float32(tor[1]) * DEG2RAD_F32))
        range_lists.append(torsion_ranges(tor[2], use_tor_lib))
        rot_lists.append([int(v) for v in tor[3]])
    frags = split_torsions_into_frags(obj.get("root_atoms", []), list(zip(axis, rot_lists)))
    pairs = intra_pairs(n, p1213, p14, frags)
    lig = make_ligand(name, coords, types, dihe, axis, rot_lists, range_lists, pairs, ff, q, obj)
    lig.pairs_1213, lig.pairs_14, lig.frags = p1213, p14, frags
    lig.bias_start, lig.bias = bias_sites(atoms)
    return lig


def bias_sites(atoms: list):
    """Per-atom bias sites, the ninth field of an atom line (rapidjson_parser.cpp:113-124):
    a list of [x, y, z, V, r2] per site, in atom order (as task.cu:157-168 lays them out).
    Returns (bias_start, bias) or (None, None) when no atom has a site."""
    start = np.zeros(len(atoms) + 1, dtype=np.int32)
    rows = []
    for ia, a in enumerate(atoms):
        sites = a[8] if len(a) > 8 else []
        for b in sites:
            if len(b) < 5:
                raise ValueError(f"bias site of atom {ia} has {len(b)} values, expected 5")
            rows.append([float(v) for v in b[:5]])
        start[ia + 1] = len(rows)
    if not rows:
        return None, None
    return start, np.asarray(rows, dtype=np.float32).reshape(-1, 5)


def make_ligand(name, coords, types, dihedrals, axis, rot_lists, range_lists, pairs, ff_types=None, charges=None,
                json_obj=None) -> Ligand:
    nt = len(axis)
    rot_start = np.zeros(nt + 1, dtype=np.int32)
    range_start = np.zeros(nt + 1, dtype=np.int32)
    for t in range(nt):
        rot_start[t + 1] = rot_start[t] + len(rot_lists[t])
        range_start[t + 1] = range_start[t] + len(range_lists[t])
    rot_atoms = np.asarray([a for r in rot_lists for a in r], dtype=np.int32)
    ranges = np.asarray([x for r in range_lists for x in r], dtype=np.float32).reshape(-1, 2)
    return Ligand(
        name=name,
        coords=np.ascontiguousarray(coords, dtype=np.float32),
        types=np.ascontiguousarray(types, dtype=np.int32),
        dihedrals=np.asarray(dihedrals, dtype=np.float32).reshape(nt),
        axis=np.asarray(axis, dtype=np.int32).reshape(nt, 2),
        rot_start=rot_start,
        rot_atoms=rot_atoms,
        range_start=range_start,
        ranges=ranges,
        pairs=np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2),
        ff_types=ff_types,
        charges=charges,
        json_obj=json_obj,
    )


def receptor_from_json(atoms: list, box: Box) -> Receptor:
    """Receptor atoms [x, y, z, xs, ff, q] with H dropped and the box+8 A crop (rapidjson_parser.cpp:25-52)."""
    grown = box.grown(VINA_CUTOFF)
    keep_xyz, keep_t, keep_ff, keep_q = [], [], [], []
    for a in atoms:
        t = int(a[3])
        if t == XS_H:
            continue
        xyz = np.asarray(a[0:3], dtype=np.float32)
        if np.all(xyz >= grown.lo) and np.all(xyz <= grown.hi):
            keep_xyz.append(xyz)
            keep_t.append(t)
            keep_ff.append(int(a[4]))
            keep_q.append(float(a[5]))
    return Receptor(
        xyz=np.asarray(keep_xyz, dtype=np.float32).reshape(-1, 3),
        types=np.asarray(keep_t, dtype=np.int32),
        ff_types=np.asarray(keep_ff, dtype=np.int32),
        charges=np.asarray(keep_q, dtype=np.float32),
    )
