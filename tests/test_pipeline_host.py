"""Host-side logic of the drop-in pipeline (no GPU): presets, grouping, output format."""
import json

import numpy as np
import pytest

from unidock_engine.api.python import pipeline as P


def test_signature_and_defaults():
    import inspect
    sig = inspect.signature(P.DockingPipeline.__init__)
    kw = [p for p in sig.parameters.values() if p.name != "self"]
    assert all(p.kind == p.KEYWORD_ONLY for p in kw)
    names = [p.name for p in kw]
    # api/python/pipeline.cpp:167-189
    assert names == ["output_dir", "center_x", "center_y", "center_z", "size_x", "size_y", "size_z", "task",
                     "search_mode", "exhaustiveness", "randomize", "mc_steps", "opt_steps", "refine_steps",
                     "num_pose", "rmsd_limit", "energy_range", "seed", "bias", "bias_k", "constraint_docking",
                     "use_tor_lib", "energy_decomp", "gpu_device_id", "name_json", "max_gpu_mem"]
    d = {p.name: p.default for p in kw}
    assert d["search_mode"] == "balance" and d["exhaustiveness"] == 512 and d["mc_steps"] == 20
    assert d["opt_steps"] == -1 and d["seed"] == 12345 and d["name_json"] == "from_python_obj"


def test_search_mode_presets():
    for mode, (ex, mc) in {"fast": (128, 20), "balance": (256, 30), "detail": (512, 40)}.items():
        p = dict(search_mode=mode, exhaustiveness=1, mc_steps=1, opt_steps=5)
        P.apply_search_mode(p)
        assert (p["exhaustiveness"], p["mc_steps"], p["opt_steps"]) == (ex, mc, -1)
    p = dict(search_mode="free", exhaustiveness=3, mc_steps=4, opt_steps=5)
    P.apply_search_mode(p)
    assert (p["exhaustiveness"], p["mc_steps"], p["opt_steps"]) == (3, 4, 5)
    with pytest.raises(ValueError):
        P.apply_search_mode(dict(search_mode="turbo"))


def test_task_presets():
    p = dict(task="score", exhaustiveness=512, num_pose=10)
    P.apply_task(p)
    assert p["exhaustiveness"] == 1 and p["mc_steps"] == 0 and p["num_pose"] == 1 and not p["randomize"]
    p = dict(task="randomize", exhaustiveness=64)
    P.apply_task(p)
    assert p["num_pose"] == 64 and p["rmsd_limit"] == 0.0
    with pytest.raises(ValueError):
        P.apply_task(dict(task="benchmark"))


def test_size_groups():
    class L:
        def __init__(self, n, t):
            self.natom, self.ntors = n, t
    ligs = [L(30, 4), L(30, 9), L(90, 10), L(200, 3), L(40, 8), L(161, 0)]
    g = P.group_ligands(ligs)
    assert g == [[0, 4], [1], [2], [], [3, 5]]


@pytest.mark.parametrize("v,s", [(1.5, "1.5"), (2.0, "2.0"), (-7.123456, "-7.123"), (0.0999, "0.099"),
                                 (1e-7, "0.0"), (float("nan"), "1000000.0"), (3e9, "1000000.0"),
                                 (-2e7, "-1000000.0"), (np.float32(0.1), "0.1"), (123.4567, "123.456")])
def test_float_format(v, s):
    """rapidjson Writer with SetMaxDecimalPlaces(3) (json.cpp:167-169) and safe_val (json.cpp:80-85)."""
    assert P.fmt_float(v) == s


def test_output_document():
    class R:
        name = "lig\"1"
        energies = [np.asarray([-5.0, -6.0, 0.5, -6.5, 0.0, 1.0, 0.0], np.float32)]
        coords = [np.asarray([1.0, 2.5, -3.25], np.float32)]
        dihedrals = [np.asarray([], np.float32)]
    txt = P.dumps_poses([R()])
    import json
    doc = json.loads(txt)
    assert list(doc) == ['lig"1'] and doc['lig"1'][0]["coords"] == [1.0, 2.5, -3.25]
    assert " " not in txt and txt.endswith("}")


def test_write_retries(tmp_path):
    path = str(tmp_path / "x.json")
    P.write_json_with_retry(path, '{"a":1}')
    assert open(path).read() == '{"a":1}'
    with pytest.raises(OSError):
        P.write_json_with_retry(str(tmp_path / "missing" / "x.json"), "{}", retries=2, wait=0.01)


def test_unsupported_options_fail_loudly():
    """apply_bias (core.cpp:86-97) rejects an unknown bias type before any GPU work."""
    pipe = P.DockingPipeline(center_x=0, center_y=0, center_z=0, size_x=10, size_y=10, size_z=10, bias="hbond")
    with pytest.raises(RuntimeError, match="Core pipeline failed") as ei:
        pipe.run()
    assert "Not supported bias" in str(ei.value.__cause__)


def test_decomp_member_in_output_document():
    """json.cpp:134-147: "decomp" follows "dihedrals", one [5] row per atom."""
    from unidock_engine.engine import LigandResult
    r = LigandResult("m", coords=[np.zeros(6, np.float32)], dihedrals=[np.zeros(0, np.float32)],
                     energies=[np.arange(7, dtype=np.float32)],
                     decomp=[np.asarray([[-0.5, -0.25, 0.125, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0, 0.0]], np.float32)])
    doc = json.loads(P.dumps_poses([r]))
    pose = doc["m"][0]
    assert list(pose) == ["energy", "coords", "dihedrals", "decomp"]
    assert pose["decomp"] == [[-0.5, -0.25, 0.125, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0, 0.0]]


def _native_params(**kw):
    p = dict(P.DEFAULTS, center_x=0.0, center_y=0.0, center_z=0.0, size_x=30.0, size_y=30.0, size_z=30.0)
    p.update(kw)
    return p


def test_native_run_rejects_bad_options_and_input(tmp_path):
    """ud2_run_json's checks that come before any device work: options, JSON syntax, topology."""
    from unidock_engine import _native as N
    out = str(tmp_path)
    good_atom = [0.0, 0.0, 0.0, 2, 3, 0.0, [], []]
    cases = [
        (dict(bias="turbo"), "{}", "Not supported bias: turbo"),
        (dict(search_mode="warp"), "{}", "Not supported search_mode: warp"),
        (dict(task="dock"), "{}", "Not supported task: dock"),
        (dict(), '{"a": [1, 2,]}', "JSON parse error at offset 12: unexpected character"),
        (dict(), '{"a": 1} x', "trailing characters"),
        (dict(), '{"lig": {"atoms": [[0.0, 1.0]]}}', "missing array element"),
        (dict(), json.dumps({"lig": {"torsions": []}}), "ligand lig: no atoms"),
        (dict(), json.dumps({"lig": {"atoms": [good_atom + [[[1.0, 2.0]]]]}}), "bias site of atom 0"),
        (dict(use_tor_lib=True), json.dumps({"lig": {"atoms": [good_atom] * 4, "torsions": [
            [[0, 1, 2, 3], 10.0, [[-30.0, -60.0]], [2, 3]]]}}), "wrong torsion range"),
    ]
    for kw, text, msg in cases:
        with pytest.raises(N.Ud2Error, match=msg):
            N.run_json(_native_params(output_dir=out, name_json="x", **kw), text)
    # no ligands: nothing written, no error
    assert N.run_json(_native_params(output_dir=out, name_json="x"), '{"receptor": [], "score": ["vina"]}') == 0
    assert not list(tmp_path.iterdir())


def test_memory_budget_batches():
    """Batches of a size group: at most 1,000 ligands, estimated device memory under the budget
    (run_screening, screening.cpp:168-196); a budget below one ligand is an error."""
    from unidock_engine import synthetic as S
    ligs = [s.ligand for s in S.config1_ligands(12)]
    group = list(range(len(ligs)))
    one = [P.BatchSize().add(l).bytes(4096) for l in ligs]
    assert all(1 << 19 < b < 4 << 20 for b in one)  # 0.5-4 MB per ligand at 4,096 chains
    batches = list(P.memory_batches(ligs, group, 4096, 8))
    assert [i for b in batches for i in b] == group and len(batches) > 2
    for b in batches:
        est = P.BatchSize()
        for i in b:
            est = est.add(ligs[i])
        assert est.bytes(4096) // P.MIB < 8
    assert list(P.memory_batches(ligs, group, 4096, 10 ** 6)) == [group]
    big = list(range(2500))
    assert [len(b) for b in P.memory_batches([ligs[0]] * 2500, big, 1, 10 ** 6)] == [1000, 1000, 500]
    with pytest.raises(RuntimeError, match="Not enough GPU Memory"):
        list(P.memory_batches(ligs, group, 4096, 1))
