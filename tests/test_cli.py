"""The ud2 command line (unidock/unidock_engine/bin/unidock/main.cpp:54-195) without a GPU:
arguments, the config template, config reading and the failure exits."""
import os
import subprocess
import sys

import pytest
import yaml

from unidock_engine import cli

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# the values of the reference's feature-test config E/test/feature/bias_hbond/ud2.yaml, with its tags
BIAS_HBOND_YAML = """
Advanced:
  seed: !!int 1234567
  exhaustiveness: !!int 512
  randomize: !!bool true
  mc_steps: !!int 40
  opt_steps: !!int -1
  refine_steps: !!int 0
  rmsd_limit: !!float 1.0
  num_pose: !!int 10
  energy_range: !!float 10
  tor_lib: !!bool false
  bias: !!str pos
  bias_k: 1.0
Hardware:
  ncpu: !!int 10
  gpu_device_id: !!int 0
  max_gpu_memory: !!int 0
Settings:
  task: !!str screen
  search_mode: !!str free
  constraint_docking: !!bool false
  center_x: !!float 2
  center_y: !!float 60
  center_z: !!float 10
  size_x: !!float 30
  size_y: !!float 30.0
  size_z: !!float 30
Outputs:
  dir: !!str ./
Inputs:
  json: !!str ../../data/bias_hbond/input.json
"""


def write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def test_help_and_missing_config(capsys):
    assert cli.main(["--help"]) == 0
    assert "Usage: ud2 [options] [config_path]" in capsys.readouterr().out
    assert cli.main([]) == 1
    assert "Missing argument: config file path" in capsys.readouterr().out
    assert cli.main(["--log"]) == 1
    assert "--log requires a log path" in capsys.readouterr().out


def test_launcher_runs_as_a_program():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "uni-dock2_amd", "bin", "ud2"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "--dump_config" in r.stdout


def test_dump_config_template(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    assert cli.main(["--dump_config"]) == 0
    doc = yaml.safe_load(open(tmp_path / "default.yaml"))
    assert list(doc) == ["Settings", "Advanced", "Hardware", "Outputs", "Inputs"]
    assert doc["Settings"]["task"] == "screen" and doc["Settings"]["size_x"] == 30.0
    assert doc["Advanced"]["exhaustiveness"] == 512 and doc["Advanced"]["mc_steps"] == 20
    assert doc["Advanced"]["seed"] == 12345 and doc["Advanced"]["refine_steps"] == 5
    assert doc["Hardware"] == {"gpu_device_id": 0, "max_gpu_memory": 0}
    assert doc["Outputs"]["output_dir"] == "./res_ud2" and doc["Inputs"]["json"] is None
    assert cli.main(["--dump_config", str(tmp_path / "x.yaml")]) == 0
    assert (tmp_path / "x.yaml").read_text() == (tmp_path / "default.yaml").read_text()
    # read back by the CLI's reader, the template's text values are the defaults ("no" stays a string)
    p, fp_json = cli.read_config(str(tmp_path / "x.yaml"))
    assert p["bias"] == "no" and p["bias_k"] == pytest.approx(0.1) and p["randomize"] is True
    assert fp_json == ""


def test_read_config_of_the_bias_feature(tmp_path):
    p, fp_json = cli.read_config(write(tmp_path / "ud2.yaml", BIAS_HBOND_YAML))
    assert fp_json == "../../data/bias_hbond/input.json"
    assert p == dict(num_pose=10, rmsd_limit=1.0, energy_range=10.0, seed=1234567, exhaustiveness=512,
                     randomize=True, mc_steps=40, opt_steps=-1, refine_steps=0, use_tor_lib=False, bias="pos",
                     bias_k=1.0, energy_decomp=False, task="screen", search_mode="free", constraint_docking=False,
                     center_x=2.0, center_y=60.0, center_z=10.0, size_x=30.0, size_y=30.0, size_z=30.0,
                     gpu_device_id=0, max_gpu_mem=0, output_dir="./")


def test_missing_items_take_defaults(tmp_path, caplog):
    path = write(tmp_path / "c.yaml", "Settings:\n  center_x: 1.5\nInputs:\n  json: a.json\n")
    with caplog.at_level("WARNING", logger="unidock_engine"):
        p, fp_json = cli.read_config(path)
    assert p["seed"] == 12345 and p["mc_steps"] == 20 and p["search_mode"] == "balance"
    assert p["output_dir"] == "" and fp_json == "a.json"
    assert (p["center_x"], p["center_y"], p["size_z"]) == (1.5, 0.0, 0.0)
    assert "Config item Advanced.seed doesn't exist" in caplog.text


@pytest.mark.parametrize("item", ["seed: 1.5", "seed: abc", "randomize: 2", "exhaustiveness: [1, 2]"])
def test_bad_values_are_fatal(tmp_path, item):
    with pytest.raises(cli.ConfigError):
        cli.read_config(write(tmp_path / "c.yaml", f"Advanced:\n  {item}\n"))
    assert cli.main(["--log", str(tmp_path / "u.log"),
                     write(tmp_path / "c.yaml", f"Advanced:\n  {item}\nInputs:\n  json: a.json\n")]) == 1
    assert "Failed to read config item Advanced." in (tmp_path / "u.log").read_text()


def test_failure_exits(tmp_path, monkeypatch, capsys):
    monkeypatch.chdir(tmp_path)
    assert cli.main([write(tmp_path / "a.yaml", "Advanced:\n  seed: 1\n")]) == 1  # empty json path
    assert cli.main([write(tmp_path / "b.yaml", "Inputs:\n  json: lig.txt\n")]) == 1  # not a .json name
    assert cli.main([write(tmp_path / "c.yaml", "Inputs:\n  json: nothere.json\n")]) == 1  # unreadable
    write(tmp_path / "in.json", "{\"receptor\": []}")
    # an unknown bias fails inside the pipeline: "Docking pipeline failed", exit 1
    assert cli.main([write(tmp_path / "d.yaml", "Advanced:\n  bias: turbo\nInputs:\n  json: in.json\n")]) == 1
    assert "Docking pipeline failed" in (tmp_path / "ud.log").read_text()
    # main.cpp skips the argument after the config path, so a third one is unexpected
    assert cli.main([str(tmp_path / "d.yaml"), "--log", "x.log"]) == 1
    assert "unexpected argument: x.log" in capsys.readouterr().out
