"""Parity at the benchmark configurations' own sizes.

- The plain monochromatic chain (N = 2^20, 32 slices, 0.1 m hop) against the
  complex128 oracle, which runs it in seconds.
- The polychromatic batch (64 energies x 2^24) and a served phase-stepping
  request at 2^28, both beyond the oracle's reach: the reference is this
  library's plain complex128 form, itself checked against the oracle at
  smaller N (test_gpu_parity.py).
- The Talbot-Lau chain with its sphere phantom: against the oracle at 2^18
  and against the plain complex128 form at 2^26 (294 steps).
Pattern: big-wave's end-to-end check, big-wave/test.py:853-960."""
import math
import shutil

import numpy as np
import pytest

import simcases
from oracle import bigwave as bw

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ravesim
    c = ravesim.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def fringe_phase_error(rows, ref):
    """Largest per-pixel difference of the phase-stepping fringe phase
    -atan2(sum_k I_k sin(2 pi k / n), sum_k I_k cos(2 pi k / n)) over the
    pixels whose first harmonic exceeds 1e-3 of the mean intensity."""
    n = rows.shape[0]
    w = np.exp(2j * np.pi * np.arange(n) / n)[:, None]
    a, b = (w * rows).sum(0), (w * ref).sum(0)
    ok = np.abs(b) > 1e-3 * np.abs(ref).mean() * n
    assert ok.sum() > 10
    return float(np.max(np.abs(np.angle(a[ok] * np.conj(b[ok])))))


_MONO_REF = {}


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-10), ("c8", 1e-4)])
def test_mono_chain_at_2_20_matches_oracle(ctx, tmp_path, dtype, tol):
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    sd = write_simdir(tmp_path / "sim", *simcases.mono_chain(tmp_path, N=2 ** 20, slices=32, det_size=1e-3,
                                                            dtype=dtype))
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    if "ref" not in _MONO_REF:
        _MONO_REF["ref"] = bw.run_simdir(sd, 0, dtype=np.complex128)
    ref = _MONO_REF["ref"]
    assert got.shape == ref.shape == (1, 1000)
    err = rel_l2(got, ref)
    print(f"mono chain 2^20 x 33 steps, {dtype}: rel-L2 vs oracle {err:.2e}")
    assert err < tol


def test_polychromatic_batch_at_64_x_2_24(ctx):
    import ravesim
    from ravesim.core import BatchMember, run_batch
    N, slices, h = 2 ** 24, 32, 30e-6
    members = []
    for e_ev, w, db in simcases.poly_members(64, N, slices):
        g = ravesim.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=h, nr_steps=slices, db_a=db)
        members.append(BatchMember(wl=ravesim.energy_to_wavelength(e_ev), gratings=[g],
                                   cutoff_angles=[math.pi / 2] * 2, weight=w, plane_wave=True))
    kw = dict(det_size=1e-3, det_px=1e-6, det_py=1.0, z_detector=h + 0.1)
    fused = run_batch(ctx, N, 2e-9, members, dtype=np.complex64, **kw)
    plain = run_batch(ctx, N, 2e-9, members, dtype=np.complex128, plain=True, max_resident=1, **kw)
    err = rel_l2(fused, plain)
    print(f"64 x 2^24 fused c64 vs plain c128: rel-L2 {err:.2e}")
    assert fused.shape == (1, 1000) and err < 1e-4


def test_served_request_at_2_28_matches_plain_c128():
    """A served phase-stepping sweep at N = 2^28, the second-largest served
    grid size (8 G2 positions, 650 steps), in complex64 against the same
    chain run position by position in the plain complex128 form."""
    import ravesim
    from ravesim.core import BatchMember, run_batch
    from ravesim.serve import Server, phase_sweep
    N = 2 ** 28
    req = phase_sweep(N, [46000.0], n_steps=8, slices=64)
    c = ravesim.Context(0)
    srv = Server(c)
    got = srv.result(srv.submit(req), timeout_s=600).rows
    srv.close()
    p = 4.2e-6
    members = []
    for k in range(8):
        grs = [ravesim.Grating(**{**g.__dict__, "x_positions": [k * p / 8] if i == 2 else [0.0]})
               for i, g in enumerate(req.gratings[0])]
        members.append(BatchMember(wl=ravesim.energy_to_wavelength(46000.0), gratings=grs,
                                   cutoff_angles=list(req.cutoff_angles), plane_wave=True, out_index=k))
    ref = run_batch(c, N, req.dx, members, det_size=req.det_size, det_px=req.det_px, det_py=1.0,
                    z_detector=req.z_detector, n_outputs=8, dtype=np.complex128, plain=True, max_resident=1)
    c.close()
    err, dphi = rel_l2(got, ref), fringe_phase_error(got, ref)
    print(f"served 2^28 c64 vs plain c128: rel-L2 {err:.2e}, fringe phase {dphi:.2e} rad")
    assert got.shape == ref.shape == (8, 1000)
    assert err < 1e-4 and dphi < 1e-3


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-10), ("c8", 1e-4)])
def test_talbot_lau_with_sphere_matches_oracle(ctx, tmp_path, dtype, tol):
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    sd = write_simdir(tmp_path / "sim", *simcases.talbot_lau_phantom(tmp_path, N=2 ** 18, slices=8,
                                                                    det_size=2e-4, dtype=dtype, pz=10e-6))
    shutil.copy(tmp_path / "sphere.npy", sd / "sphere.npy")
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    assert rel_l2(got, ref) < tol


def test_talbot_lau_with_sphere_at_2_26(ctx, tmp_path):
    """The 2^26 Talbot-Lau chain with its optional sphere phantom at full
    size: 3 x 64 slices, 2 hops and 100 sphere rows (1 um z pixels),
    complex64 against the plain complex128 form."""
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    sd = write_simdir(tmp_path / "sim", *simcases.talbot_lau_phantom(tmp_path, N=2 ** 26, dtype="c8"))
    shutil.copy(tmp_path / "sphere.npy", sd / "sphere.npy")
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = run_single_simulation(sd, 0, tmp_path, ctx=ctx, dtype="c16", plain=True)
    err = rel_l2(got, ref)
    print(f"Talbot-Lau + sphere 2^26 c8 vs plain c16: rel-L2 {err:.2e}")
    assert got.shape == ref.shape == (1, 1000) and err < 1e-4
