"""Serving form: continuous batching of requests of different sizes.

Checks that a request's detector rows match the oracle, and that they are
bit-for-bit the same whether the request is served alone or while other
requests (other N, other energies) share its launches."""
import time

import numpy as np
import pytest

from oracle import bigwave as bw

pytestmark = pytest.mark.gpu


def oracle_sweep(req):
    """Weighted incoherent sum over the request's energies, each a phase-stepped
    oracle run (big-wave/wavesim.py:84-298) of the same setup."""
    total = None
    for e, w, grs in zip(req.energies_ev, req.weights or [1.0] * len(req.energies_ev), req.gratings):
        wl = bw.convert_energy_wavelength(e)
        p = bw.SimParams(N=req.N, dx=req.dx, z_detector=req.z_detector, detector_size=req.det_size,
                         detector_pixel_size_x=req.det_px, detector_pixel_size_y=req.det_py, wl=wl,
                         chunk_size=req.N)
        els = [bw.Grating(pitch=g.pitch, dc=tuple(g.dc), z_start=g.z_start, thickness=g.thickness,
                          nr_steps=g.nr_steps, x_positions=np.array(g.x_positions, dtype=float),
                          substrate_thickness=g.substrate_thickness, db_a=g.db_a, db_b=g.db_b,
                          db_substrate=g.db_substrate) for g in grs]
        rows = np.array(bw.run_simulation(p, bw.VectorSource(np.ones(req.N, complex), 0.0), els,
                                          list(req.cutoff_angles)))
        total = w * rows if total is None else total + w * rows
    return total


@pytest.fixture(scope="module")
def server():
    import ravesim
    from ravesim.serve import Server
    ctx = ravesim.Context(0)
    s = Server(ctx)
    yield s
    s.close()
    ctx.close()


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 1e-4)])
def test_served_sweep_matches_oracle(server, dtype, tol):
    from ravesim.serve import phase_sweep
    req = phase_sweep(2 ** 16, [40000.0, 52000.0], n_steps=4, slices=4, dtype=dtype, weights=[0.3, 0.7])
    got = server.result(server.submit(req)).rows
    ref = oracle_sweep(req)
    assert got.shape == ref.shape == (4, 65)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


def test_alone_and_in_the_mix_are_identical(server):
    from ravesim.serve import phase_sweep
    a = phase_sweep(2 ** 18, [46000.0, 48000.0, 50000.0], n_steps=8, slices=6)
    alone = server.result(server.submit(a)).rows
    others = [phase_sweep(2 ** 16, [30000.0 + 1000 * i], n_steps=8, slices=6) for i in range(3)]
    others += [phase_sweep(2 ** 20, [60000.0], n_steps=8, slices=3)]
    ids = [server.submit(others[0]), server.submit(others[3])]
    time.sleep(0.01)
    ida = server.submit(a)
    ids += [server.submit(o) for o in others[1:3]]
    mixed = server.result(ida).rows
    for i in ids:
        server.result(i)
    st = server.stats()
    assert st["jobs_per_launch"] > 1.0  # launches were shared
    assert np.array_equal(alone, mixed)


def _free_device_bytes():
    import ctypes as C
    free = C.c_size_t()
    total = C.c_size_t()
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemGetInfo(C.byref(free), C.byref(total))
    return free.value


def test_admission_with_small_budget_completes():
    """A budget that holds one job at a time still serves every request:
    the second job waits for the first one's buffers and reuses them."""
    import ravesim
    from ravesim.serve import Server, phase_sweep
    ctx = ravesim.Context(0)
    one_job = 2 * 2 ** 22 * 8 + 8 * 500 * 8  # field + snapshot + detector rows
    s = Server(ctx, reserve_bytes=float(_free_device_bytes() - int(1.5 * one_job)))
    reqs = [phase_sweep(2 ** 22, [46000.0 + 500 * i], n_steps=2, slices=2) for i in range(3)]
    ids = [s.submit(r) for r in reqs]
    outs = [s.result(i, timeout_s=120).rows for i in ids]
    npx = int(reqs[0].det_size // reqs[0].det_px)
    assert all(np.isfinite(o).all() and o.shape == (2, npx) for o in outs)
    assert s.stats()["max_resident_jobs"] == 1
    s.close()
    ctx.close()


def test_job_that_can_never_fit_fails_its_request():
    """With a reserve above the device's memory no job fits even alone: the
    request fails with a memory message instead of waiting forever, later
    requests are refused the same way, and the server shuts down."""
    import ravesim
    from ravesim._lib import RsError
    from ravesim.serve import Server, phase_sweep
    ctx = ravesim.Context(0)
    s = Server(ctx, reserve_bytes=float(_free_device_bytes()) * 4)
    t0 = time.perf_counter()
    ids = [s.submit(phase_sweep(2 ** 16, [46000.0], n_steps=2, slices=2)) for _ in range(2)]
    for i in ids:
        with pytest.raises(RsError, match="RS_ERR_NOMEM"):
            s.result(i, timeout_s=60)
    s.close()
    assert time.perf_counter() - t0 < 30
    ctx.close()


def test_context_is_reserved_while_a_server_lives():
    import ravesim
    from ravesim._lib import RsError
    from ravesim.serve import Server
    ctx = ravesim.Context(0)
    f = ravesim.Field(ctx, 2 ** 12, 2e-9)
    s = Server(ctx)
    for call in (lambda: f.propagate(2.7e-11, 1e-3), lambda: ravesim.Field(ctx, 2 ** 12, 2e-9),
                 lambda: ctx.autotune(2 ** 12), lambda: Server(ctx)):
        with pytest.raises(RsError, match="reserved by a server"):
            call()
    s.close()
    f.propagate(2.7e-11, 1e-3)  # usable again
    f.close()
    ctx.close()


@pytest.mark.parametrize("seed", range(4))
def test_random_request_mix_matches_oracle(server, seed):
    """A seeded mix of requests submitted together: N = 2^12..2^15, 1-3
    energies with random weights, 2-4 phase steps, 2-4 slices per grating,
    complex64 or complex128; every result against its oracle sweep."""
    from ravesim.serve import phase_sweep
    rng = np.random.default_rng(300 + seed)
    reqs = []
    for _ in range(int(rng.integers(3, 7))):
        dt = np.complex128 if rng.random() < 0.5 else np.complex64
        ne = int(rng.integers(1, 4))
        reqs.append(phase_sweep(int(2 ** rng.integers(12, 16)), list(rng.uniform(20e3, 60e3, ne)),
                                n_steps=int(rng.integers(2, 5)), slices=int(rng.integers(2, 5)),
                                det_size=2e-6, dtype=dt, weights=list(rng.uniform(0.2, 1.0, ne))))
    ids = [server.submit(r) for r in reqs]
    for r, i in zip(reqs, ids):
        got = server.result(i).rows
        ref = oracle_sweep(r)
        tol = 1e-10 if r.dtype == np.complex128 else 1e-4
        assert got.shape == ref.shape
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol, (r.N, r.dtype)
