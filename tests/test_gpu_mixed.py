"""Mixed batches: simulations of different grid sizes in one call
(rs_batch_run_mixed).  Each simulation's row against the complex128 oracle,
bit for bit the row a single-size run_batch gives it, and the whole run
captured as one hipGraph over both lanes equal to the eager run on every
replay (with eager work and synchronisations between the replays)."""
import math

import numpy as np
import pytest

from oracle import bigwave as bw

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def mixed_case(seed, n_sims=7, lg=(12, 16)):
    """Seeded uneven mix: N = 2^lg[0]..2^(lg[1]-1), energy 20-60 keV, pitch
    0.4-1.5 um, duty cycle 0.4-0.6, 1-3 point sources each (summed
    incoherently), one Si grating of 2-3 slices, cutoff from the detector
    geometry; detector half the field."""
    import ravesim
    from ravesim.core import BatchMember, BatchSim
    rng = np.random.default_rng(seed)
    sims, refs = [], []
    dx = 1e-9
    for _ in range(n_sims):
        N = int(2 ** rng.integers(*lg))
        e = float(rng.uniform(20e3, 60e3))
        wl = ravesim.energy_to_wavelength(e)
        pitch = float(rng.uniform(0.4e-6, 1.5e-6))
        dc = float(rng.uniform(0.4, 0.6))
        steps = int(rng.integers(2, 4))
        db = complex(rng.uniform(2e-7, 2e-6), rng.uniform(0, 5e-8))
        g = ravesim.Grating(pitch=pitch, dc=(dc, dc), z_start=0.01, thickness=4e-6, nr_steps=steps, db_a=db)
        og = bw.Grating(pitch=pitch, dc=(dc, dc), z_start=0.01, thickness=4e-6, nr_steps=steps,
                        x_positions=np.array([0.0]), substrate_thickness=0.0, db_a=db)
        ang = sorted(float(v) for v in rng.choice([1e-3, 2e-3, math.pi / 2], 2))
        det, z_det = N * dx / 2, 0.03
        ms, ref = [], None
        p = bw.SimParams(N=N, dx=dx, z_detector=z_det, detector_size=det, detector_pixel_size_x=2e-7,
                         detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
        for x in rng.uniform(-5e-8, 5e-8, int(rng.integers(1, 4))):
            w = float(rng.uniform(0.3, 1.0))
            ms.append(BatchMember(wl=wl, gratings=[g], cutoff_angles=ang, weight=w, source_x=float(x), source_z=-0.2))
            row = w * bw.run_simulation(p, bw.PointSource(x=float(x), z=-0.2), [og], ang)[0]
            ref = row if ref is None else ref + row
        sims.append(BatchSim(N=N, dx=dx, members=ms, det_size=det, det_px=2e-7, z_detector=z_det))
        refs.append(ref)
    return sims, refs


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-10), (np.complex64, 1e-4)])
def test_mixed_sizes_match_oracle_and_single_size_runs(seed, dtype, tol):
    import ravesim
    from ravesim.core import run_batch, run_mixed
    sims, refs = mixed_case(40 + seed)
    c = ravesim.Context(0)
    rows = run_mixed(c, sims, dtype=dtype)
    assert len({s.N for s in sims}) >= 2
    for s, row, ref in zip(sims, rows, refs):
        assert rel_l2(row, ref) < tol, s.N
    # the same simulations as single-size batches, one per part of the mixed
    # run: grid size, detector and structure (here the slice count)
    by_n = {}
    for i, s in enumerate(sims):
        by_n.setdefault((s.N, s.det_size, s.members[0].gratings[0].nr_steps), []).append(i)
    for (N, det, _), idx in by_n.items():
        ms = []
        for r, i in enumerate(idx):
            for m in sims[i].members:
                m.out_index = r
                ms.append(m)
        single = run_batch(c, N, sims[idx[0]].dx, ms, n_outputs=len(idx), dtype=dtype, det_size=det,
                           det_px=2e-7, det_py=1.0, z_detector=sims[idx[0]].z_detector)
        for r, i in enumerate(idx):
            assert np.array_equal(single[r], rows[i])
    c.close()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_mixed_graph_replays_match_eager(dtype):
    import ravesim
    from ravesim.core import run_mixed
    sims, _ = mixed_case(7, n_sims=9, lg=(14, 19))
    c = ravesim.Context(0)
    eager = run_mixed(c, sims, dtype=dtype)
    for _ in range(3):
        got = run_mixed(c, sims, dtype=dtype, use_graph=True)
        c.synchronize()
        for a, b in zip(got, eager):
            assert np.array_equal(a, b)
        for a, b in zip(run_mixed(c, sims, dtype=dtype), eager):
            assert np.array_equal(a, b)
    c.close()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_graph_replay_takes_new_parameters(dtype):
    """A second graph run of the same shape replays the cached graph with its
    own argument records: other energies, pitches, duty cycles and source
    positions give that run's eager rows, not the first run's; a run of
    another shape captures anew."""
    import ravesim
    from ravesim.core import run_mixed
    a, _ = mixed_case(21, n_sims=5, lg=(13, 16))
    b, _ = mixed_case(21, n_sims=5, lg=(13, 16))
    rng = np.random.default_rng(5)
    for s in b:  # same structure, new physics
        for m in s.members:
            m.wl *= float(rng.uniform(0.8, 1.2))
            m.source_x += 3e-8
            g = m.gratings[0]
            m.gratings = [ravesim.Grating(pitch=g.pitch * 1.1, dc=(0.47, 0.47), z_start=g.z_start,
                                          thickness=g.thickness, nr_steps=g.nr_steps, db_a=g.db_a * 1.3)]
    other, _ = mixed_case(22, n_sims=4, lg=(12, 14))
    c = ravesim.Context(0)
    ea, eb, eo = (run_mixed(c, x, dtype=dtype) for x in (a, b, other))
    for sims, ref in ((a, ea), (b, eb), (a, ea), (other, eo), (b, eb)):
        got = run_mixed(c, sims, dtype=dtype, use_graph=True)
        for g, r in zip(got, ref):
            assert np.array_equal(g, r)
    assert not all(np.array_equal(x, y) for x, y in zip(ea, eb))
    c.close()


def test_mixed_batch_errors():
    import ravesim
    from ravesim._lib import RsError
    from ravesim.core import BatchMember, BatchSim, run_mixed
    c = ravesim.Context(0)
    g = ravesim.Grating(pitch=1e-6, dc=(0.5, 0.5), z_start=0.01, thickness=4e-6, nr_steps=2, db_a=1e-6 + 0j,
                        x_positions=[0.0, 1e-7])
    m = BatchMember(wl=3e-11, gratings=[g], cutoff_angles=[math.pi / 2] * 2, source_z=-0.1)
    with pytest.raises(RsError, match="phase stepped"):
        run_mixed(c, [BatchSim(N=2 ** 12, dx=1e-9, members=[m], det_size=1e-6, det_px=1e-7, z_detector=0.02)])
    with pytest.raises(RsError, match="power of two"):
        m.gratings[0].x_positions = [0.0]
        run_mixed(c, [BatchSim(N=3000, dx=1e-9, members=[m], det_size=1e-6, det_px=1e-7, z_detector=0.02)])
    c.close()
