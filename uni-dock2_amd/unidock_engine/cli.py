"""The ``ud2`` command line: ``ud2 [options] [config_path]``.

Follows the reference's CLI (unidock/unidock_engine/bin/unidock/main.cpp:54-195):

* ``--version``, ``--help``, ``--dump_config [path]`` (default ``default.yaml``, the template of
  bin/unidock/main.h:31-114), ``--log <path>`` (default ``ud.log`` in the working directory).
* The config is YAML with sections Advanced / Settings / Hardware / Outputs / Inputs, read key by
  key as main.cpp:120-165 reads them: a missing key (or one with an empty value) logs a warning
  and takes its default (CoreInputDefaults, src/screening/core.h:9-36; the box and the paths
  default to 0 / ""), a value of the wrong type is fatal (exit 1). Keys are main.cpp's: ``Advanced.tor_lib`` and
  ``Outputs.dir`` (the dumped template names them ``use_tor_lib`` and ``output_dir``, which
  main.cpp does not read; so does this reader).
* ``Inputs.json`` is an engine-JSON file (receptor + ligands); its stem names the outputs,
  ``<Outputs.dir>/<stem>_<batch>.json``; a name not ending in ``.json`` is an error (exit 1).
* The docking run is the same core pipeline the Python ``DockingPipeline`` drives; a failure
  logs "Docking pipeline failed" and exits 1.

As in main.cpp:89-99, the argument right after the config path is skipped.
"""
from __future__ import annotations

import logging
import os
import sys

import yaml


log = logging.getLogger("unidock_engine")

HELP = """
Usage: ud2 [options] [config_path]

Options:
  --version             Show version information
  --help                Show this help message
  --dump_config [path]  Dump default config to file (default: default.yaml)
  --log <path>          Specify log file path, only take effects when [config_path] is given
"""

SIGN = r"""
     _   _ ____ ____
    | | | |  _ \___ \
    | | | | | | |__) |
    | |_| | |_| / __/
     \___/|____/_____|   MI355X engine
"""

# CoreInputDefaults (src/screening/core.h:9-36) with DockParam's values (src/model/model.h:151-168)
DEFAULTS = dict(
    bias="no", bias_k=0.1, constraint_docking=False, energy_range=10.0, exhaustiveness=512, gpu_device_id=0,
    mc_steps=20, max_gpu_memory=0, num_pose=10, name_json="from_python_obj", opt_steps=-1,
    output_dir="./res_ud2", randomize=True, refine_steps=5, rmsd_limit=1.0, search_mode="balance", seed=12345,
    task="screen", energy_decomp=False, use_tor_lib=False,
)

# CoreInputDocs (src/screening/core.h:41-67)
DOCS = dict(
    bias="(str) Bias type. Options: no, pos (position), align",
    bias_k="(float) Bias scaling coefficient",
    constraint_docking="(bool) Enable constraint docking (disable translation & orientation DOFs)",
    energy_range="(float) Energy range for output poses (kcal/mol)",
    exhaustiveness="(int) Number of independent MC runs (roughly proportional to time)",
    gpu_device_id="(int) GPU device ID to use",
    max_gpu_memory="(int) Max GPU memory in MB (0 for all available)",
    mc_steps="(int) Monte Carlo random walk steps per run",
    num_pose="(int) Maximum number of output poses per ligand",
    opt_steps="(int) Optimization steps after each MC step (-1 for auto)",
    output_dir="(str) Output directory for docking results",
    randomize="(bool) Whether to randomize input pose before global search",
    refine_steps="(int) Refinement steps after clustering",
    rmsd_limit="(float) RMSD threshold for pose clustering (Angstrom)",
    search_mode="(str) Search mode. Options: fast, balance, detail, free",
    seed="(int) Random seed for reproducibility",
    task="(str) Docking task type. Options: screen, score, mc",
    energy_decomp="(bool) Output per-atom inter energy decomposition (gauss1, gauss2, repulsion, hydrophobic, hbond)",
    use_tor_lib="(bool) Use torsion angle library",
)

# (section, key, type, default) in the order main.cpp:120-165 reads them; the CoreInput field each
# fills is the key unless named in FIELD
ITEMS = (
    ("Advanced", "num_pose", int, DEFAULTS["num_pose"]),
    ("Advanced", "rmsd_limit", float, DEFAULTS["rmsd_limit"]),
    ("Advanced", "energy_range", float, DEFAULTS["energy_range"]),
    ("Advanced", "seed", int, DEFAULTS["seed"]),
    ("Advanced", "exhaustiveness", int, DEFAULTS["exhaustiveness"]),
    ("Advanced", "randomize", bool, DEFAULTS["randomize"]),
    ("Advanced", "mc_steps", int, DEFAULTS["mc_steps"]),
    ("Advanced", "opt_steps", int, DEFAULTS["opt_steps"]),
    ("Advanced", "refine_steps", int, DEFAULTS["refine_steps"]),
    ("Advanced", "tor_lib", bool, DEFAULTS["use_tor_lib"]),
    ("Advanced", "bias", str, DEFAULTS["bias"]),
    ("Advanced", "bias_k", float, DEFAULTS["bias_k"]),
    ("Advanced", "energy_decomp", bool, DEFAULTS["energy_decomp"]),
    ("Settings", "task", str, DEFAULTS["task"]),
    ("Settings", "search_mode", str, DEFAULTS["search_mode"]),
    ("Settings", "constraint_docking", bool, DEFAULTS["constraint_docking"]),
    ("Settings", "center_x", float, 0.0),
    ("Settings", "center_y", float, 0.0),
    ("Settings", "center_z", float, 0.0),
    ("Settings", "size_x", float, 0.0),
    ("Settings", "size_y", float, 0.0),
    ("Settings", "size_z", float, 0.0),
    ("Hardware", "gpu_device_id", int, DEFAULTS["gpu_device_id"]),
    ("Hardware", "max_gpu_memory", int, DEFAULTS["max_gpu_memory"]),
    ("Outputs", "dir", str, ""),
    ("Inputs", "json", str, ""),
)
FIELD = {"tor_lib": "use_tor_lib", "dir": "output_dir", "max_gpu_memory": "max_gpu_mem"}

_TRUE = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON", "y", "Y"}
_FALSE = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF", "n", "N"}


class ConfigError(Exception):
    """A config value that cannot be read as its type (fatal, main.cpp:44-47)."""


def _convert(node, typ, where: str):
    """YAML::Node::as<T>() on a scalar node's text (yaml-cpp keeps the text; an explicit tag such as
    ``!!int`` does not change it): an int reads integer text only, a float any number, a bool
    YAML's true/false words, a string the text itself (so ``bias: no`` is the string "no")."""
    if not isinstance(node, yaml.ScalarNode):
        raise ConfigError(f"Failed to read config item {where}: not a scalar")
    text = node.value
    if typ is str:
        return text
    if typ is bool:
        if text in _TRUE or text in _FALSE:
            return text in _TRUE
    elif typ is int:
        try:
            return int(text, 0)
        except ValueError:
            pass
    elif typ is float:
        try:
            return float(text)
        except ValueError:
            pass
    raise ConfigError(f"Failed to read config item {where}: bad conversion of {text!r} to {typ.__name__}")


def _mapping(node) -> dict:
    """A YAML mapping node as {key text: value node}."""
    if not isinstance(node, yaml.MappingNode):
        return {}
    return {k.value: v for k, v in node.value if isinstance(k, yaml.ScalarNode)}


def read_config(path: str):
    """Parse a ``ud2`` YAML config (main.cpp:118-165). Returns (params, json_path): params holds
    DockingPipeline's keyword arguments (box centre and size included) except name_json.
    Raises ConfigError."""
    with open(path) as f:
        root = yaml.compose(f, Loader=yaml.SafeLoader)
    if root is not None and not isinstance(root, yaml.MappingNode):
        raise ConfigError(f"{path}: the config is not a mapping")
    config = _mapping(root)
    vals = {}
    for sec, key, typ, default in ITEMS:
        node = _mapping(config.get(sec)).get(key)
        if node is None or (isinstance(node, yaml.ScalarNode) and node.tag.endswith(":null")):
            log.warning("Config item %s.%s doesn't exist, using default value", sec, key)
            v = default
        else:
            v = _convert(node, typ, f"{sec}.{key}")
        vals[FIELD.get(key, key)] = v
    fp_json = vals.pop("json")
    return vals, fp_json


def _yaml_scalar(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, float):
        return f"{v:.3f}".rstrip("0").rstrip(".") if v != int(v) else f"{v:.1f}"
    return str(v)


def dump_config_template(path: str) -> None:
    """The default config template (bin/unidock/main.h:31-114): Settings, Advanced, Hardware,
    Outputs and Inputs with default values and one comment per key."""
    rows = [("Settings", [(k, DEFAULTS[k], DOCS[k]) for k in ("task", "search_mode", "constraint_docking")] + [
        ("center_x", 0.0, "float: X coordinate of box center (Angstrom) - REQUIRED"),
        ("center_y", 0.0, "float: Y coordinate of box center (Angstrom) - REQUIRED"),
        ("center_z", 0.0, "float: Z coordinate of box center (Angstrom) - REQUIRED"),
        ("size_x", 30.0, "float: Box size along X axis (Angstrom) - REQUIRED"),
        ("size_y", 30.0, "float: Box size along Y axis (Angstrom) - REQUIRED"),
        ("size_z", 30.0, "float: Box size along Z axis (Angstrom) - REQUIRED")]),
        ("Advanced", [(k, DEFAULTS[k], DOCS[k]) for k in (
            "seed", "exhaustiveness", "randomize", "mc_steps", "opt_steps", "refine_steps", "rmsd_limit",
            "num_pose", "energy_range", "bias", "bias_k", "use_tor_lib", "energy_decomp")]),
        ("Hardware", [(k, DEFAULTS[k], DOCS[k]) for k in ("gpu_device_id", "max_gpu_memory")]),
        ("Outputs", [("output_dir", DEFAULTS["output_dir"], DOCS["output_dir"])]),
        ("Inputs", [("json", None, "str: Input JSON file - REQUIRED")])]
    out = ["# Uni-Dock2 Configuration Template", "# Generated with default values from CoreInputDefaults", ""]
    for sec, items in rows:
        out.append(f"{sec}:")
        for k, v, doc in items:
            val = "" if v is None else " " + _yaml_scalar(v)
            out.append(f"  {k}:{val}  # {doc}")
    try:
        with open(path, "w") as f:
            f.write("\n".join(out) + "\n")
    except OSError:
        sys.stderr.write(f"Failed to create config template file: {path}\n")
        raise SystemExit(1)


_handlers: list = []


def _close_logger() -> None:
    for h in _handlers:
        log.removeHandler(h)
        h.close()
    _handlers.clear()


def _init_logger(fp_log: str) -> None:
    """init_logger: console and file, info level."""
    _close_logger()
    log.setLevel(logging.INFO)
    fmt = logging.Formatter("[%(asctime)s] [%(levelname)s] %(message)s")
    for h in (logging.StreamHandler(sys.stdout), logging.FileHandler(fp_log, mode="a")):
        h.setFormatter(fmt)
        log.addHandler(h)
        _handlers.append(h)


def _version() -> str:
    from . import _native as N
    return N.version()


def main(argv=None) -> int:
    """main(), bin/unidock/main.cpp:54-195. Returns the process exit code."""
    argv = list(sys.argv[1:] if argv is None else argv)
    fp_log, fp_config = "ud.log", ""
    print(SIGN)
    i = 0
    while i < len(argv):
        arg = argv[i]
        if arg == "--version":
            print(f"UD2 Engine Version: {_version()}")
            return 0
        if arg == "--help":
            print(HELP, end="")
            return 0
        if arg == "--dump_config":
            fp_dump = "default.yaml"
            if i + 1 < len(argv) and not argv[i + 1].startswith("-"):
                i += 1
                fp_dump = argv[i]
            dump_config_template(fp_dump)
            print(f"Dump config template file: {fp_dump}")
            return 0
        if arg == "--log":
            if i + 1 < len(argv):
                fp_log = argv[i + 1]
                i += 1
            else:
                print(HELP, end="")
                print("--log requires a log path")
                return 1
        elif not fp_config:
            fp_config = arg
            i += 1  # main.cpp:92 skips the argument after the config path
        else:
            print(HELP, end="")
            print(f"After config path is parsed as {fp_config}, unexpected argument: {arg}")
            return 1
        i += 1
    if not fp_config:
        print(HELP, end="")
        print("Missing argument: config file path")
        return 1

    print(f"UD2 Engine Version: {_version()}")
    _init_logger(fp_log)
    try:
        return _run_config(fp_config)
    finally:
        _close_logger()


def _run_config(fp_config: str) -> int:
    """main.cpp:116-192: read the config and the input JSON, run the pipeline."""
    log.info("Using config file: %s", fp_config)
    try:
        params, fp_json = read_config(fp_config)
    except (ConfigError, OSError, yaml.YAMLError) as e:
        log.critical("%s", e)
        return 1
    if not fp_json:
        log.critical("Empty json file path")
        return 1
    name = os.path.basename(fp_json)
    if len(name) < 5 or not name.endswith(".json"):
        log.error("Wrong JSON file name: %s", name)
        return 1
    params["name_json"] = name[:-5]
    try:
        with open(fp_json, encoding="utf-8") as f:
            text = f.read()
    except (OSError, ValueError) as e:  # read_ud_from_json, src/format/json.cpp:24-41
        log.critical("Failed to read JSON file %s: %s", fp_json, e)
        return 1

    from .api.python.pipeline import run_native
    try:
        run_native(params, text)
    except Exception as e:  # core_pipeline returns 1 (src/screening/core.cpp:218-222)
        log.critical("Core pipeline failed: %s", e)
        log.critical("Docking pipeline failed")
        return 1
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
