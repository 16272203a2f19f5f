"""HIP path against the committed oracle fixtures (tests/golden/*.npz) of the
reduced benchmark configurations."""
import math

import numpy as np
import pytest

import simcases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ravesim
    c = ravesim.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _simdir(tmp_path, case):
    from ravesim.simdir import write_simdir
    cfg, subs, ang = case
    return write_simdir(tmp_path / "sim", cfg, subs, ang)


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-11), ("c8", 1e-4)])
def test_mono_fixture(ctx, golden_dir, tmp_path, dtype, tol):
    from ravesim.multisim import run_single_simulation
    fx = np.load(golden_dir / "mono_n16.npz")
    sd = _simdir(tmp_path, simcases.mono_chain(tmp_path, N=int(fx["N"]), slices=int(fx["slices"]),
                                         det_size=float(fx["det_size"]), dtype=dtype))
    got = run_single_simulation(sd, 0, ctx=ctx)
    assert rel_l2(got, fx["rows"]) < tol


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-11), ("c8", 1e-4)])
def test_talbot_lau_phase_stepping_fixture(ctx, golden_dir, tmp_path, dtype, tol):
    from ravesim.multisim import run_single_simulation
    fx = np.load(golden_dir / "talbot_lau_n18_steps8.npz")
    sd = _simdir(tmp_path, simcases.talbot_lau(tmp_path, N=int(fx["N"]), slices=int(fx["slices"]),
                                         det_size=float(fx["det_size"]), dtype=dtype,
                                         g2_steps=int(fx["g2_steps"])))
    got = run_single_simulation(sd, 0, ctx=ctx)
    assert got.shape == fx["rows"].shape
    assert rel_l2(got, fx["rows"]) < tol


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 1e-4)])
def test_polychromatic_batch_fixture(ctx, golden_dir, dtype, tol):
    import ravesim
    from ravesim.core import BatchMember, run_batch
    fx = np.load(golden_dir / "poly_n16_e8.npz")
    N, slices = int(fx["N"]), int(fx["slices"])
    h = 30e-6
    members = []
    for e_ev, w, db in simcases.poly_members(int(fx["n_energies"]), N, slices):
        g = ravesim.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=h, nr_steps=slices,
                            db_a=db)
        members.append(BatchMember(wl=ravesim.energy_to_wavelength(e_ev), gratings=[g],
                                   cutoff_angles=[math.pi / 2] * 2, weight=w, plane_wave=True))
    got = run_batch(ctx, N, 2e-9, members, det_size=1e-4, det_px=1e-6, det_py=1.0,
                    z_detector=h + 0.1, dtype=dtype)
    assert rel_l2(got, fx["rows"]) < tol
    # same members split over smaller resident groups give the same sum
    got2 = run_batch(ctx, N, 2e-9, members, det_size=1e-4, det_px=1e-6, det_py=1.0,
                     z_detector=h + 0.1, dtype=dtype, max_resident=3)
    assert rel_l2(got2, got) < 1e-14
