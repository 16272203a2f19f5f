"""Host-side checks that need no GPU: the C-ABI library loads and exports
every declared symbol, the C++ sim-dir parser / validator, the exact detector
pixel ranges, and the sim-dir producer's setup rules against the oracle's."""
import json
import math
import re
from pathlib import Path

import numpy as np
import pytest
import yaml

from oracle import bigwave as bw
from simcases import mono_chain, talbot_lau

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    from ravesim import _lib
    header = (ROOT / "include" / "ravesim.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(rs_\w+)\(", header, re.M))
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(_lib.lib, name), name
    assert _lib.lib.rs_version() == _lib.VERSION
    assert re.search(r"#define RS_VERSION (\d+)", header).group(1) == str(_lib.VERSION)


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror in ravesim._lib has the size and the field offsets
    of its declaration in include/ravesim.h, compiled by the host C compiler
    (a field missing from either side fails to compile or to match)."""
    import ctypes
    import subprocess
    from ravesim import _lib
    header = (ROOT / "include" / "ravesim.h").read_text()
    structs = re.findall(r"^typedef struct (rs_\w+) \{", header, re.M)
    assert len(structs) >= 10
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ravesim.h"', "int main(void) {"]
    for s in structs:
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for name, _ in getattr(_lib, s)._fields_:
            lines.append(f'printf("{s}.{name} %zu\\n", offsetof({s}, {name}));')
    lines.append("return 0; }")
    (tmp_path / "layout.c").write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-I", str(ROOT / "include"), str(tmp_path / "layout.c"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.splitlines())
    for s in structs:
        cls = getattr(_lib, s)
        assert int(got[s]) == ctypes.sizeof(cls), s
        for name, _ in cls._fields_:
            assert int(got[f"{s}.{name}"]) == getattr(cls, name).offset, f"{s}.{name}"


def test_failure_without_device_is_an_error_code():
    import ravesim
    from ravesim import _lib
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(_lib.RsError) as e:
        ravesim.Context(0)
    assert e.value.code in (_lib.RS_ERR_DEVICE, _lib.RS_ERR_INVALID)


def _write(tmp_path, case):
    from ravesim.simdir import write_simdir
    cfg, subs, ang = case
    return write_simdir(tmp_path / "sim", cfg, subs, ang)


def test_validate_mono_and_talbot_lau(tmp_path):
    from ravesim.multisim import validate_simdir
    sd = _write(tmp_path, mono_chain(tmp_path, N=2 ** 16, slices=32, det_size=1e-4))
    N, bits, steps, npx, nel, props = validate_simdir(sd, 0)
    assert (N, bits, steps, npx, nel, props) == (2 ** 16, 16, 1, 100, 1, 33)
    sd2 = _write(tmp_path / "b", talbot_lau(tmp_path, N=2 ** 18, slices=64, dtype="c8", g2_steps=8))
    N, bits, steps, npx, nel, props = validate_simdir(sd2, 0)
    assert (N, bits, steps, nel, props) == (2 ** 18, 8, 8, 3, 194)


def test_parser_reads_ruamel_style_yaml(tmp_path):
    """Block sequences indented under their key and compact nested lists, as
    big-wave's ruamel dumper writes them (indent mapping=2, sequence=4, offset=2)."""
    from ravesim.multisim import validate_simdir
    sd = tmp_path / "sim"
    (sd / "00000000").mkdir(parents=True)
    np.save(tmp_path / "ones.npy", np.ones(1024, dtype=np.complex128))
    (sd / "config.yaml").write_text(f"""# comment
sim_params:
  N: 1024
  dx: 8e-10   # trailing comment
  z_detector: 1.77
  detector_size: 2e-7
  detector_pixel_size_x: 1e-8
  detector_pixel_size_y: 1e-8
  chunk_size: 10469376
use_disk_vector: False
save_final_u_vectors: false
dtype: c8
elements:
  - type: grating
    pitch: 4.2e-06
    dc:
      - 0.5
      - 0.5
    z_start: 0.1
    thickness: 0.00014
    nr_steps: 2
    x_positions: [0.0, 1.0e-6]
    substrate_thickness: 0.00023
    mat_a:
      - Si
      - 2.34
    mat_b: ["Au", 19.32]
    mat_substrate: null
""")
    (sd / "computed.yaml").write_text("cutoff_angles:\n  - 0.1\n  - 0.2\n")
    (sd / "00000000" / "subconfig.yaml").write_text(f"""energy: 46000.0
source:
  type: vector
  input_path: '{tmp_path / "ones.npy"}'
  z: 0.1
deltabeta_table:
  - - - Si
      - 2.34
    - - 2.2888e-07
      - 2.5847e-10
  - - - Au
      - 19.32
    - - 1.4981e-06
      - 3.6174e-08
""")
    N, bits, steps, npx, nel, props = validate_simdir(sd, 0)
    assert (N, bits, steps, npx, nel) == (1024, 8, 2, int(2e-7 // 1e-8), 1)  # 19: Python floor division
    assert props == 2 + 1 + 1  # two slices, substrate, final hop


@pytest.mark.parametrize("mutation,needle", [
    (lambda c, s: c["sim_params"].pop("dx"), "dx"),
    (lambda c, s: c.__setitem__("dtype", "c32"), "Unknown dtype"),
    (lambda c, s: c["elements"][0].__setitem__("z_start", -1.0), "Overlapping"),
    (lambda c, s: s[0]["deltabeta_table"].clear(), "not found"),
    (lambda c, s: c["elements"][0].__setitem__("type", "lens"), "not supported"),
    (lambda c, s: c["sim_params"].__setitem__("N", 1000), "power of two"),
    (lambda c, s: s[0]["deltabeta_table"][0].__setitem__(1, [float("nan"), 0.0]), "not finite"),
])
def test_config_errors_are_reported(tmp_path, mutation, needle):
    from ravesim import _lib
    from ravesim.multisim import validate_simdir
    cfg, subs, ang = mono_chain(tmp_path, N=2 ** 12, slices=2, det_size=4e-6)
    mutation(cfg, subs)
    sd = _write(tmp_path, (cfg, subs, ang))
    with pytest.raises(_lib.RsError) as e:
        validate_simdir(sd, 0)
    assert needle in str(e.value)
    assert e.value.code in (_lib.RS_ERR_CONFIG, _lib.RS_ERR_UNSUPPORTED)


def test_missing_simdir_is_io_error(tmp_path):
    from ravesim import _lib
    from ravesim.multisim import validate_simdir
    with pytest.raises(_lib.RsError) as e:
        validate_simdir(tmp_path / "nothing", 0)
    assert e.value.code == _lib.RS_ERR_IO


@pytest.mark.parametrize("N,dx,size,px", [
    (8, 0.5, 3.0, 1.0), (118, 0.2, 10.0, 0.3), (118, 0.2, 2.0, 0.6), (4096, 2e-9, 4e-6, 1e-7),
    (4096, 3.1e-10, 2.5e-7, 1e-8), (10000, 1e-9, 5e-6, 7e-8), (2 ** 16, 2e-9, 1e-4, 1e-6)])
def test_detector_bounds_reproduce_histogram(N, dx, size, px):
    """Each sample's pixel from the host ranges equals np.histogram's bin."""
    from ravesim.core import detector_bounds
    b = detector_bounds(N, dx, size, px)
    o = int(size // px)
    assert len(b) == o + 1 and np.all(np.diff(b) >= 0)
    pos = (np.arange(N) - N / 2) * dx
    rng = (-(o / 2) * px, (o / 2 + 1) * px)
    # integer weights make every pixel sum exact, so equality is bin equality
    rs = np.random.default_rng(0)
    w = rs.integers(1, 1000, N).astype(np.float64)
    exp = np.histogram(pos, bins=o + 1, range=rng, weights=w)[0][:-1]
    got = np.array([w[b[k]:b[k + 1]].sum() for k in range(o)])
    assert np.array_equal(got, exp)
    cnt = np.histogram(pos, bins=o + 1, range=rng)[0][:-1]
    assert np.array_equal(np.diff(b), cnt)


def test_setup_rules_match_oracle(tmp_path, golden_dir):
    """Product sim-dir producer: same RNG draws and cutoff angles as the oracle."""
    from ravesim.simdir import setup_simulation
    import yaml
    d = json.loads((golden_dir / "bigwave_snapshot.json").read_text())
    cfg = d["known_config"]
    sd = setup_simulation(cfg, tmp_path / "sim", lambda m, r, e: (1e-7, 1e-10))
    comp = yaml.safe_load((sd / "computed.yaml").read_text())
    sub = yaml.safe_load((sd / "00000000" / "subconfig.yaml").read_text())
    ms = cfg["multisource"]
    E, X = bw.draw_point_sources(ms["seed"], 1, ms["energy_range"], ms["x_range"])
    assert sub["energy"] == float(E[0]) and sub["source"]["x"] == float(X[0])
    els = bw.elements_from_config(cfg["elements"], [(("Si", 2.34), 0j), (("Au", 19.32), 0j)])
    ang, _ = bw.compute_cutoff_angles(cfg["sim_params"]["detector_size"], cfg["sim_params"]["dx"],
                                      ms["energy_range"], ms["z"], [e.z_start for e in els],
                                      [e.get_thickness() for e in els], cfg["sim_params"]["z_detector"],
                                      1e-6)
    assert np.allclose(comp["cutoff_angles"], ang, rtol=1e-15, atol=0)


@pytest.mark.parametrize("mutation,needle", [
    (lambda c, g: g.__setitem__(0, np.full((2, 3), 3, dtype=np.int8)), "more materials"),
    (lambda c, g: c["elements"][1].__setitem__("z_start", 0.0), "z_start"),
    (lambda c, g: c["elements"][1].__setitem__("pixel_size_z", 0.0), "pixel sizes"),
    (lambda c, g: g.__setitem__(0, np.zeros((2, 3), dtype=np.float64)), "int8"),
])
def test_sample_config_errors(tmp_path, mutation, needle):
    """Sample elements are validated like big-wave's Sample.check_valid."""
    from ravesim import _lib
    from ravesim.multisim import validate_simdir
    from ravesim.simdir import write_simdir
    from simcases import sample_case
    cfg, subs, ang = sample_case(tmp_path)
    grid = [np.load(tmp_path / "sample_grid.npy")]
    mutation(cfg, grid)
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    np.save(sd / "sample_grid.npy", grid[0])
    with pytest.raises(_lib.RsError) as e:
        validate_simdir(sd, 0)
    assert needle in str(e.value)


def test_sample_config_parses(tmp_path):
    """A valid sample sim dir: one propagation per grid row, phase stepped."""
    from ravesim.multisim import validate_simdir
    from ravesim.simdir import write_simdir
    from simcases import sample_case
    cfg, subs, ang = sample_case(tmp_path)
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    np.save(sd / "sample_grid.npy", np.load(tmp_path / "sample_grid.npy"))
    info = validate_simdir(sd, 0)
    # N, c16, 2 phase steps, 40 px, 2 elements, 1 source prop + 4 slices + hop + 6 rows + hop
    assert list(info) == [2 ** 16, 16, 2, 40, 2, 13]


def test_dba_config_parses(tmp_path):
    """fast-wave's dba_grating element (config_parsing.cpp:107-137): 6 slices
    + substrate, phase stepped over two positions."""
    from ravesim.multisim import validate_simdir
    from ravesim.simdir import write_simdir
    from simcases import dba_case
    cfg, subs, ang = dba_case()
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    assert list(validate_simdir(sd, 0))[:4] == [2 ** 16, 16, 2, 40]
    cfg["elements"][0]["thickness"] = 0.0
    sd = write_simdir(tmp_path / "bad", cfg, subs, ang)
    from ravesim import _lib
    with pytest.raises(_lib.RsError) as e:
        validate_simdir(sd, 0)
    assert "thickness" in str(e.value)


def test_setup_simulation_with_sample_and_dba(tmp_path):
    """The sim-dir producer collects sample materials, takes a sample's
    thickness from its grid and copies the grid (big-wave/multisim.py:489-499)."""
    from ravesim.multisim import validate_simdir
    from ravesim.simdir import setup_simulation
    from simcases import dba_case, sample_case
    cfg, _, _ = sample_case(tmp_path)
    dcfg, _, _ = dba_case()
    cfg["elements"].insert(1, dict(dcfg["elements"][0], z_start=0.011))
    cfg["multisource"] = {"type": "points", "nr_source_points": 2, "energy_range": [29000.0, 31000.0],
                          "x_range": [-1e-7, 1e-7], "z": -0.3, "seed": 3}
    cfg["sim_params"]["N"] = 2 ** 20
    sd = setup_simulation(cfg, tmp_path / "sim", lambda m, r, e: (1e-7, 1e-10), config_dir=tmp_path)
    assert (sd / "sample_grid.npy").exists()
    sub = yaml.safe_load((sd / "00000001" / "subconfig.yaml").read_text())
    assert [m[0] for m, _ in sub["deltabeta_table"]] == ["Si", "Au"]
    comp = yaml.safe_load((sd / "computed.yaml").read_text())
    assert len(comp["cutoff_angles"]) == 4
    assert list(validate_simdir(sd, 1))[:2] == [2 ** 20, 16]


FASTWAVE = Path(__file__).resolve().parent.parent / "rave-sim_amd" / "ravesim" / "fastwave"


@pytest.mark.skipif(not FASTWAVE.exists(), reason="fastwave is not built")
def test_fastwave_cli_arguments():
    """fast-wave's command line (fast-wave/main.cpp:25-57): -h prints the help
    and exits 0; a missing source index or sim dir is an error with exit 1."""
    import subprocess
    r = subprocess.run([str(FASTWAVE), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--source_idx" in r.stdout and "--history_dz" in r.stdout
    assert "--tune" in r.stdout
    r = subprocess.run([str(FASTWAVE), "x", "-s", "0", "--tune"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "missing value for tune" in r.stderr
    r = subprocess.run([str(FASTWAVE), "/nonexistent"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Source index is required" in r.stderr
    r = subprocess.run([str(FASTWAVE), "-s", "0"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Simulation directory is required" in r.stderr
    r = subprocess.run([str(FASTWAVE), "x", "-s", "0", "--history_dz", "-1"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "history_dz" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 exits 2 with a message, before
    touching a device."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_malformed_yaml_is_a_config_error(tmp_path):
    """A config.yaml that is not YAML is reported, not crashed on."""
    from ravesim import _lib
    from ravesim.multisim import validate_simdir
    sd = _write(tmp_path, mono_chain(tmp_path, N=2 ** 12, slices=2, det_size=4e-6))
    (sd / "config.yaml").write_text("sim_params: {N: 4096, dx: [1e-9\nelements: - : -\n\t]\n")
    with pytest.raises(_lib.RsError) as e:
        validate_simdir(sd, 0)
    assert e.value.code == _lib.RS_ERR_CONFIG and str(e.value)


def test_null_handles_are_rejected():
    """Every entry point checks its handles and pointers and returns
    RS_ERR_INVALID with a message, without touching a device."""
    import ctypes as C
    from ravesim import _lib
    lib = _lib.lib
    inv = _lib.RS_ERR_INVALID
    checks = [
        lambda: lib.rs_propagate(None, 2e-11, 1e-3, 1e30),
        lambda: lib.rs_field_upload(None, None),
        lambda: lib.rs_field_download(None, None),
        lambda: lib.rs_field_copy(None, None),
        lambda: lib.rs_point_source(None, 2e-11, 0.0, 1.0),
        lambda: lib.rs_apply_grating(None, 2e-11, None, 0, 1e30),
        lambda: lib.rs_apply_dba_grating(None, 2e-11, None, 0, 1e30),
        lambda: lib.rs_apply_sample(None, 2e-11, None, 0, 1e30),
        lambda: lib.rs_scale(None, 1.0, 0.0),
        lambda: lib.rs_square_and_downsample(None, None, 1.0, None, 0),
        lambda: lib.rs_run_simdir(None, b"x", 0, None),
        lambda: lib.rs_batch_run(None, None, None, 0),
        lambda: lib.rs_field_create(None, None, _lib.RS_C8, 0, None),
        lambda: lib.rs_ctx_synchronize(None),
        lambda: lib.rs_set_timing(None, 1),
        lambda: lib.rs_autotune(None, 1024, _lib.RS_C8, 1, None, None),
        lambda: lib.rs_set_split(None, 1024, _lib.RS_C8, 6, 4),
        lambda: lib.rs_get_pass_stats(None, None, 0),
        lambda: lib.rs_server_create(None, 0.0, 0, None),
        lambda: lib.rs_server_submit(None, None, None),
        lambda: lib.rs_server_poll(None, 0, None, None, 0, None, None),
        lambda: lib.rs_server_wait(None, 0, 0.0, None),
        lambda: lib.rs_server_stats(None, None),
    ]
    for k, call in enumerate(checks):
        assert call() == inv, k
        assert lib.rs_last_error().decode()


def test_yaml_anchors_and_aliases(tmp_path):
    """Objects shared in a config dict are written by PyYAML as anchors and
    aliases (a material used by several gratings, a repeated element); they
    parse to the same configuration as the written-out form."""
    from ravesim.multisim import validate_simdir
    from ravesim.simdir import write_simdir
    from simcases import grating
    si = ["Si", 2.34]
    g = grating(1e-6, 0.01, 4e-6, 3, si, substrate=1e-6, mat_s=si)
    def cfg(els):
        return {"sim_params": {"N": 4096, "dx": 1e-9, "z_detector": 0.05, "detector_size": 1e-6,
                               "detector_pixel_size_x": 1e-7, "detector_pixel_size_y": 1.0, "chunk_size": 4096},
                "use_disk_vector": False, "save_final_u_vectors": False, "dtype": "c16", "elements": els}
    subs = [{"energy": 30000.0, "source": {"type": "point", "x": 0.0, "z": -1.0},
             "deltabeta_table": [[si, [5.1e-7, 1.6e-10]]]}]
    g2 = dict(g, z_start=0.02)
    shared = write_simdir(tmp_path / "a", cfg([g, g2]), subs, [0.5] * 3)
    text = (shared / "config.yaml").read_text()
    assert "&id" in text and "*id" in text
    import copy
    plain = write_simdir(tmp_path / "b", copy.deepcopy(cfg([g, g2])) | {"elements": [
        {k: (list(v) if isinstance(v, list) else v) for k, v in e.items()} for e in (g, g2)]},
        [{"energy": 30000.0, "source": {"type": "point", "x": 0.0, "z": -1.0},
          "deltabeta_table": [[["Si", 2.34], [5.1e-7, 1.6e-10]]]}], [0.5] * 3)
    assert "&id" not in (plain / "config.yaml").read_text()
    assert list(validate_simdir(shared, 0)) == list(validate_simdir(plain, 0))
    # written by hand: an anchored block mapping in a sequence, a flow
    # mapping with an anchor, aliases of a list inside later elements, and an
    # alias of a block mapping
    top = text[:text.index("elements:")]
    (shared / "config.yaml").write_text(top + """elements:
- &first
  type: grating
  pitch: 1.0e-06
  dc: [0.5, 0.5]
  z_start: 0.001
  thickness: 1.0e-06
  nr_steps: 2
  x_positions: [0.0]
  substrate_thickness: 0.0
  mat_a: &si [Si, 2.34]
- &second {type: grating, pitch: 2.0e-06, dc: [0.4, 0.4], z_start: 0.01, thickness: 2.0e-06, nr_steps: 1, x_positions: [0.0], substrate_thickness: 1.0e-06, mat_a: *si, mat_substrate: *si}
- type: grating
  pitch: 1.5e-06
  dc: [0.5, 0.5]
  z_start: 0.02
  thickness: 1.0e-06
  nr_steps: 1
  x_positions: [0.0]
  substrate_thickness: 0.0
  mat_a: *si
unused_copy: *first
""")
    (shared / "computed.yaml").write_text("cutoff_angles: [0.5, 0.5, 0.5, 0.5]\n")
    assert list(validate_simdir(shared, 0))[4] == 3
