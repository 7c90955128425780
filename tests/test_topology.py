"""Input contract (JSON -> engine topology) and the synthetic workload generators.
CPU only."""
import json
import os

import numpy as np
import pytest

from unidock_engine import topology as T

HERE = os.path.dirname(os.path.abspath(__file__))


def test_json_parse_kat():
    """E/test/unit/format/test_json.cpp:9-71 on the reference's dev1/simple.json fixture."""
    doc = json.load(open(os.path.join(HERE, "golden", "simple.json")))
    box = T.Box.from_center_size((0, 0, 0), (200, 200, 200))
    rec = T.receptor_from_json(doc["receptor"], box)
    # simple.json's receptor carries no XS hydrogen, so all five atoms are kept
    assert rec.natom == 5
    ligs = [T.ligand_from_json(k, v) for k, v in doc.items() if k not in ("score", "receptor")]
    assert [l.name for l in ligs] == ["lig1", "lig2"]
    l1, l2 = ligs
    assert l1.natom == 3 and l1.ntors == 2
    assert np.allclose(l1.centroid(), [65.484336, 14.577332, 43.487335], atol=1e-4)
    assert l1.dihedrals[0] == np.float32(np.float32(177) * T.DEG2RAD_F32)
    assert l1.dihedrals[1] == np.float32(np.float32(70) * T.DEG2RAD_F32)
    assert len(l1.pairs_1213) == 8 and len(l1.pairs_14) == 9
    assert l1.npair == 1  # the reference stores the pair in both directions: 2 entries
    assert l2.natom == 2 and l2.ntors == 2
    assert np.allclose(l2.centroid(), [63.788497, 12.872499, 41.792999], atol=1e-4)
    assert l2.dihedrals[0] == np.float32(np.float32(-13.5) * T.DEG2RAD_F32)
    assert len(l2.pairs_1213) == 5 and len(l2.pairs_14) == 6
    assert l2.npair == 1


def test_receptor_crop_and_hydrogen_drop():
    box = T.Box.from_center_size((0, 0, 0), (10, 10, 10))
    atoms = [[0, 0, 0, 2, 1, 0.0], [12.9, 0, 0, 3, 1, 0.0], [13.1, 0, 0, 3, 1, 0.0], [1, 1, 1, 0, 1, 0.0]]
    rec = T.receptor_from_json(atoms, box)
    # box +- 8 A keeps x <= 13, drops the hydrogen (rapidjson_parser.cpp:34,41)
    assert rec.natom == 2 and list(rec.types) == [2, 3]


def test_torsion_library_ranges():
    """Ranges across +-180 deg are split in two (rapidjson_parser.cpp:151-173)."""
    pi = T.PI_F32
    r = T.torsion_ranges([[170.0, -170.0], [-60, 60]], use_tor_lib=True)
    assert len(r) == 3
    assert r[0][1] == pi and r[1][0] == -pi
    assert T.torsion_ranges([[170.0, -170.0]], use_tor_lib=False) == [(-pi, pi)]
    with pytest.raises(ValueError):
        T.torsion_ranges([[-10.0, -20.0]], use_tor_lib=True)


def test_fragments_split():
    """split_torsions_into_frags (rapidjson_parser.cpp:227-296) on a chain 0-1-2-3-4-5 with
    root {0, 1} and torsions (1,2) rotating {2,3,4,5} and (3,4) rotating {4,5}."""
    frags = T.split_torsions_into_frags([0, 1], [((1, 2), [2, 3, 4, 5]), ((3, 4), [4, 5])])
    fs = sorted(sorted(f) for f in frags)
    assert fs == [[0, 1, 2], [1, 2, 3, 4], [3, 4, 5]]


@pytest.mark.parametrize("seed,n,r", [(1000, 24, 5), (3000, 40, 8), (2007, 12, 15)])
def test_synthetic_ligand_geometry(seed, n, r):
    from unidock_engine import synthetic as S
    sl = S.make_ligand(seed, n, r)
    lig = sl.ligand
    assert lig.natom == n and lig.ntors <= r
    x = lig.coords.astype(np.float64)
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    atoms = sl.json_obj["atoms"]
    bonded = {(i, j) for i in range(n) for j in atoms[i][6] if d[i, j] < 1.6}
    for i, j in bonded:
        assert abs(d[i, j] - 1.5) < 1e-4
    nb = d + np.eye(n) * 10
    for i, j in bonded:
        nb[i, j] = nb[j, i] = 10
    assert nb.min() >= 2.0 - 1e-4
    # every torsion rotates the far side of its bond, "to" included, never the "from" atom
    for t in range(lig.ntors):
        frm, to = lig.axis[t]
        rot = set(lig.rot_atoms[lig.rot_start[t]:lig.rot_start[t + 1]])
        assert to in rot and frm not in rot
    # deterministic by seed
    assert np.array_equal(S.make_ligand(seed, n, r).ligand.coords, lig.coords)


def test_synthetic_receptor():
    from unidock_engine import synthetic as S
    rec = S.make_receptor(400, seed=1)
    d = np.linalg.norm(rec.xyz - np.asarray(S.POCKET_CENTER), axis=1)
    assert d.min() >= 16.0 - 1e-4 and d.max() <= 30.0 + 1e-4
    dd = np.linalg.norm(rec.xyz[:, None] - rec.xyz[None], axis=-1) + np.eye(400) * 10
    assert dd.min() >= 3.0 - 1e-4
    assert set(np.unique(rec.types)) <= {2, 3, 4, 5, 6, 10, 11, 15}


def test_bias_sites_of_the_bias_hbond_fixture():
    """E/test/data/bias_hbond/input.json: atom 6 carries the one site [1.782, 66.634, 6.337, -15, 1.0]
    (E/test/feature/bias_hbond/validate.py:2); parsed as the ninth atom field (rapidjson_parser.cpp:113-124)."""
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    doc = json.load(open(os.path.join(here, "golden", "bias_hbond_input.json")))
    lig = T.ligand_from_json("MOL_0", doc["MOL_0"])
    assert lig.natom == 41 and lig.bias_start is not None
    counts = np.diff(lig.bias_start)
    assert counts[6] == 1 and counts.sum() == 1
    assert np.allclose(lig.bias[0], [1.782, 66.634, 6.337, -15.0, 1.0])
    plain_atoms = [a[:8] for a in doc["MOL_0"]["atoms"]]  # the same ligand without its site
    assert T.ligand_from_json("m", dict(doc["MOL_0"], atoms=plain_atoms)).bias_start is None
