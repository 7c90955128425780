"""The reference's two feature tests (E/test/feature/*/run_test.sh: ``ud2 ud2.yaml`` in the case
directory, then validate.py on the output file) run through this package's ud2 command line, with
the configs' values and the reference's own input files (tests/golden)."""
import json
import os
import shutil

import numpy as np
import pytest

from unidock_engine import cli

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))

BIN_YAML = """
Advanced:
  seed: !!int 121
  exhaustiveness: !!int 512
  randomize: !!bool true
  mc_steps: !!int 40
  opt_steps: !!int -1
  refine_steps: !!int 1
  rmsd_limit: !!float 1.0
  num_pose: !!int 10
  energy_range: !!float 10
  tor_lib: !!bool false
  energy_decomp: !!bool false
Hardware:
  gpu_device_id: !!int 0
  max_gpu_memory: !!int 0
Settings:
  task: !!str screen
  search_mode: !!str free
  constraint_docking: !!bool false
  center_x: !!float -22.33497980811559
  center_y: !!float 13.31094327905649
  center_z: !!float 27.36396790165424
  size_x: !!float 30.0
  size_y: !!float 30.0
  size_z: !!float 30.0
Outputs:
  dir: !!str ./
Inputs:
  json: !!str ../../data/5S8I/5S8I_unidock2.json
"""


def case_dirs(tmp_path, data_name, files):
    """<tmp>/test/feature/<case> and <tmp>/test/data/<data_name> as in the reference's tree."""
    case = tmp_path / "test" / "feature" / "case"
    data = tmp_path / "test" / "data" / data_name
    case.mkdir(parents=True)
    data.mkdir(parents=True)
    for src, dst in files:
        shutil.copy(os.path.join(HERE, "golden", src), data / dst)
    return case


def sdf_coords(path):
    lines = open(path).read().splitlines()
    n = int(lines[3].split()[0])
    return np.array([[float(v) for v in lines[4 + i].split()[:3]] for i in range(n)])


def test_feature_bin_5s8i(gpu, tmp_path, monkeypatch):
    """E/test/feature/bin: top poses of 5S8I redocked within 2.0 A heavy-atom RMSD of the crystal
    ligand (validate.py --rmsd-limit 2.0, best over the output poses)."""
    case = case_dirs(tmp_path, "5S8I", [("5S8I_unidock2.json", "5S8I_unidock2.json"),
                                         ("5S8I_ligand.sdf", "5S8I_ligand.sdf")])
    (case / "ud2.yaml").write_text(BIN_YAML)
    monkeypatch.chdir(case)
    print("ud2 ud2.yaml: feature bin (5S8I)", flush=True)
    assert cli.main(["ud2.yaml"]) == 0
    out = json.load(open(case / "5S8I_unidock2_1.json"))
    (name, poses), = out.items()
    ref = sdf_coords(tmp_path / "test" / "data" / "5S8I" / "5S8I_ligand.sdf")
    rmsd = [float(np.sqrt(np.mean(np.sum((np.reshape(p["coords"], (-1, 3)) - ref) ** 2, axis=1)))) for p in poses]
    print(f"top-1 RMSD {rmsd[0]:.3f} A, best {min(rmsd):.3f} A over {len(rmsd)} poses", flush=True)
    assert 0 < len(poses) <= 10 and min(rmsd) <= 2.0
    assert os.path.exists(case / "ud.log")


def test_feature_bias_hbond(gpu, tmp_path, monkeypatch):
    """E/test/feature/bias_hbond: bias "pos" (k = 1); atom 6 of each of the top 5 poses within 0.2 A
    per coordinate of its bias site (1.782, 66.634, 6.337)."""
    from test_cli import BIAS_HBOND_YAML
    case = case_dirs(tmp_path, "bias_hbond", [("bias_hbond_input.json", "input.json")])
    (case / "ud2.yaml").write_text(BIAS_HBOND_YAML)
    monkeypatch.chdir(case)
    print("ud2 ud2.yaml: feature bias_hbond", flush=True)
    assert cli.main(["ud2.yaml"]) == 0
    out = json.load(open(case / "input_1.json"))
    poses = list(out.values())[0]
    assert len(poses) >= 5
    for p in poses[:5]:
        xyz = np.asarray(p["coords"][18:21])
        assert np.all(np.abs(xyz - [1.782, 66.634, 6.337]) <= 0.2), xyz
