"""Sim-dir producer: spectrum sampling, point draws, save_and_exit and the
continuation from saved u vectors, after big-wave's own tests
(big-wave/test.py:446-491, 515-573, 592-657, 710-805)."""
import numpy as np
import pytest
import yaml

DB = lambda mat, dens, energy: (5.1e-7, 1.6e-10) if mat == "Si" else (3.6e-6, 2.1e-7)  # noqa: E731


def _grating(z, steps=1):
    return {"type": "grating", "pitch": 4.2e-6, "dc": [0.5, 0.5], "z_start": z, "thickness": 140e-6,
            "nr_steps": steps, "x_positions": [0.0], "substrate_thickness": 0,
            "mat_a": ["Si", 2.34], "mat_b": ["Au", 19.32]}


def _config(elements, n=1):
    # big-wave/test.py:395-443 and 711-770
    return {"sim_params": {"N": 4096 * 4, "dx": 8e-9, "z_detector": 1.77, "detector_size": 1e-5,
                           "detector_pixel_size_x": 1e-6, "detector_pixel_size_y": 1e-6,
                           "chunk_size": 10469376},
            "use_disk_vector": False, "save_final_u_vectors": False, "dtype": "c16",
            "multisource": {"type": "points", "energy_range": [46000.0, 47000.0], "x_range": [0.0, 1e-6],
                            "z": 0.0, "nr_source_points": n, "seed": 4},
            "elements": elements}


def _load(p):
    return yaml.safe_load(p.read_text())


def test_spectrum_energies_kat():
    """big-wave/test.py:642-657: bins 0, .5, 1, 1.5 (2.0 lies outside [0, 2)),
    only 1.5 has weight, so every energy lies in [1.5, 2)."""
    from ravesim.simdir import energies_from_spectrum
    e = np.arange(5) * 0.5
    pdf = np.zeros_like(e)
    pdf[3], pdf[4] = 2, 10
    out = energies_from_spectrum(e, pdf, np.random.RandomState(123), (0, 2), 100)
    assert len(out) == 100 and np.all(out >= 1.5) and np.all(out < 2) and np.any(out >= 1.75)
    with pytest.raises(ValueError):
        energies_from_spectrum(e, pdf, np.random.RandomState(1), (1.9, 2.4), 5)


def test_point_draws_match_the_reference_values(tmp_path):
    """big-wave/test.py:604-608 lists the draws of seed 4: energies first,
    then x positions."""
    from ravesim.simdir import setup_simulation
    sd = setup_simulation(_config([_grating(0.1)], n=2), tmp_path / "sim", DB)
    pts = _load(sd / "computed.yaml")["source_points"]
    # printed there truncated to four / eight significant digits
    assert np.allclose([p["x"] for p in pts], [9.726e-07, 7.148e-07], rtol=0, atol=1e-10)
    assert np.allclose([p["energy"] for p in pts], [46967.029, 46547.232], rtol=0, atol=1e-3)
    for i, p in enumerate(pts):
        sub = _load(sd / f"{i:08d}" / "subconfig.yaml")
        assert sub["source"] == {"type": "point", "x": p["x"], "z": 0.0} and sub["energy"] == p["energy"]


def test_spectrum_setup_draws_energies_then_positions(tmp_path):
    from ravesim.simdir import setup_simulation
    cfg = _config([_grating(0.1)], n=64)
    e = 46000.0 + 250.0 * np.arange(5)
    pdf = np.array([0.0, 1.0, 0.0, 3.0, 0.0])
    np.savez(tmp_path / "spec.npz", energy=e, pdf=pdf)
    cfg["multisource"]["spectrum"] = "spec.npz"
    sd = setup_simulation(cfg, tmp_path / "sim", DB, config_dir=tmp_path)
    pts = _load(sd / "computed.yaml")["source_points"]
    en = np.array([p["energy"] for p in pts])
    assert np.all(((en >= 46250) & (en < 46500)) | ((en >= 46750) & (en < 47000)))
    assert 0.5 < np.mean(en >= 46750) < 0.95  # weight 3 : 1
    xs = np.array([p["x"] for p in pts])
    assert np.all((xs >= 0) & (xs < 1e-6))
    # the same with the spectrum passed as arrays
    sd2 = setup_simulation(_config([_grating(0.1)], n=64), tmp_path / "sim2", DB, spectrum=(e, pdf))
    assert _load(sd2 / "computed.yaml")["source_points"] == pts


def test_save_and_exit_cuts_the_setup_and_continuation_reads_its_vectors(tmp_path):
    """big-wave/test.py:710-805 without the runs (those are GPU tests)."""
    from ravesim.multisim import validate_simdir
    from ravesim.simdir import setup_simulation
    els = [_grating(0.1), {"type": "save_and_exit", "z_start": 0.25}, _grating(0.25), _grating(0.5)]
    cfg = _config(els)
    base = setup_simulation(cfg, tmp_path / "base", DB)
    adj = _load(base / "config.yaml")
    assert adj["save_final_u_vectors"] is True
    assert len(adj["elements"]) == 1 and adj["sim_params"]["z_detector"] == 0.25
    bcomp = _load(base / "computed.yaml")
    assert len(bcomp["cutoff_angles"]) == 5  # the angles of the whole setup are kept
    assert list(validate_simdir(base, 0))[:3] == [4096 * 4, 16, 1]
    cont = dict(cfg, multisource={"type": "vectors", "base_sim_dir": str(base), "input_u_index": 0},
                elements=els[2:])
    sd = setup_simulation(cont, tmp_path / "cont", DB)
    comp = _load(sd / "computed.yaml")
    assert comp["cutoff_angles"] == bcomp["cutoff_angles"][2:]
    assert comp["source_points"] == bcomp["source_points"]
    bsub = _load(base / "00000000" / "subconfig.yaml")
    sub = _load(sd / "00000000" / "subconfig.yaml")
    assert sub["source"] == {"type": "vector", "input_path": str(base / "00000000" / "u_0000.npy"),
                             "z": 0.25}
    assert sub["energy"] == bsub["energy"]
    # a base without saved vectors cannot be continued
    plain = setup_simulation(_config([_grating(0.1)]), tmp_path / "plain", DB)
    with pytest.raises(ValueError, match="save_final_u_vectors"):
        setup_simulation(dict(cont, multisource={"type": "vectors", "base_sim_dir": str(plain),
                                                 "input_u_index": 0}), tmp_path / "bad", DB)


def test_save_and_exit_must_sit_on_the_next_element(tmp_path):
    from ravesim.simdir import setup_simulation
    els = [_grating(0.1), {"type": "save_and_exit", "z_start": 0.2}, _grating(0.25)]
    with pytest.raises(ValueError, match="same z coordinate as the next"):
        setup_simulation(_config(els), tmp_path / "sim", DB)
