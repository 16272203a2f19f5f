"""Closed-form checks of the oracle's propagation and element arithmetic.

Each check is chosen so that a plausible slip (sign of the propagator,
a missing factor pi or lambda, a wrong frequency axis, a cutoff off by one,
a sample interpolation mistake) fails it."""
import numpy as np

from oracle import bigwave as bw


def _gauss_case(N=2 ** 14, dx=2e-8, s=4e-7, wl=2.7e-11, dz=0.05):
    x = (np.arange(N) - N / 2) * dx
    u0 = np.exp(-x ** 2 / (2 * s ** 2)).astype(np.complex128)
    a = 2 * np.pi ** 2 * s ** 2 - 1j * np.pi * wl * dz
    exact = (np.sqrt(2 * np.pi) * s * np.sqrt(np.pi / a) * np.exp(-np.pi ** 2 * x ** 2 / a)
             * np.exp(-2j * np.pi * dz / wl))
    return x, u0, exact, dx, wl, dz


def test_gaussian_fresnel_closed_form():
    x, u0, exact, dx, wl, dz = _gauss_case()
    f = bw.Fields(u0.copy())
    bw.propagate(f, dx, wl, dz, len(u0), 1e30)
    err = np.linalg.norm(f.u - exact) / np.linalg.norm(exact)
    # the global phase 2 pi dz/wl ~ 1e10 rad carries ~1e-6 rad of FP64 rounding
    assert err < 1e-5
    assert np.linalg.norm(np.abs(f.u) - np.abs(exact)) / np.linalg.norm(exact) < 1e-12


def test_gaussian_conjugate_convention_fails():
    # the opposite sign convention must be far from the reference one
    x, u0, exact, dx, wl, dz = _gauss_case()
    f = bw.Fields(u0.copy())
    bw.propagate(f, dx, wl, dz, len(u0), 1e30)
    assert np.linalg.norm(f.u - np.conj(exact)) / np.linalg.norm(exact) > 0.1


def test_two_hops_compose_and_preserve_norm():
    rng = np.random.default_rng(3)
    u = rng.standard_normal(4096) + 1j * rng.standard_normal(4096)
    dx, wl = 3e-9, 2e-11
    f1 = bw.Fields(u.copy())
    bw.propagate(f1, dx, wl, 0.03, 4096, 1e30)
    bw.propagate(f1, dx, wl, 0.07, 1000, 1e30)
    f2 = bw.Fields(u.copy())
    bw.propagate(f2, dx, wl, 0.10, 4096, 1e30)
    assert np.allclose(f1.u, f2.u, rtol=0, atol=1e-10 * np.abs(u).max())
    assert abs(np.linalg.norm(f2.u) / np.linalg.norm(u) - 1) < 1e-12


def test_cutoff_band_is_zero_after_propagate():
    u = np.ones(64, dtype=np.complex128)
    u[5] = 3
    dx, wl, fc = 1e-9, 2e-11, 1.7e8
    f = bw.Fields(u)
    bw.propagate(f, dx, wl, 0.01, 64, fc)
    spec = np.fft.fft(f.u)
    band = np.abs(np.fft.fftfreq(64) / dx) >= fc
    assert band.any() and np.abs(spec[band]).max() < 1e-12
    assert np.abs(spec[~band]).min() > 1e-6


def test_uniform_sample_equals_uniform_grating():
    # a sample made of one material everywhere is the same slab as a grating
    # whose two materials are that material
    N, dx, wl = 1024, 2e-9, 2.5e-11
    db = complex(3e-7, 2e-9)
    p = bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                     detector_pixel_size_y=1.0, wl=wl, chunk_size=300)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    grid = np.ones((3, 40), dtype=np.int8)
    s = bw.Sample(z_start=0.0, pixel_size_x=1e-7, pixel_size_z=2e-6, grid=grid,
                  db_list=np.array([0, db], dtype=np.complex128), x_positions=np.array([0.0]))
    g = bw.Grating(pitch=1e-6, dc=(0.5, 0.5), z_start=0.0, thickness=6e-6, nr_steps=3,
                   x_positions=np.array([0.0]), substrate_thickness=0.0, db_a=db, db_b=db)
    fs, fg = bw.Fields(u.copy()), bw.Fields(u.copy())
    # restrict to the sample's footprint: both see the same factor inside it
    bw.sample_apply(fs, s, p, 1e30, 0)
    bw.grating_apply(fg, g, p, 1e30, 0)
    # outside the footprint (|x| > 2e-6) the sample is vacuum; compare with a
    # field that is zero there so both elements act identically
    x = (np.arange(N) - N / 2) * dx
    inside = np.abs(x) < 1.5e-6
    u2 = np.where(inside, u, 0)
    fs, fg = bw.Fields(u2.copy()), bw.Fields(u2.copy())
    for r in range(3):
        fs.u *= np.where(inside, bw.material_factor(db, 2e-6, wl), 1.0)
        bw.propagate(fs, dx, wl, 2e-6, 300, 1e30)
    bw.grating_apply(fg, g, p, 1e30, 0)
    ref = bw.Fields(u2.copy())
    bw.sample_apply(ref, s, p, 1e30, 0)
    assert np.allclose(ref.u, fs.u, rtol=0, atol=1e-12 * np.abs(u).max())


def _sample_params(N=2048, dx=2e-9, wl=2.5e-11, chunk=300):
    return bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                        detector_pixel_size_y=1.0, wl=wl, chunk_size=chunk)


def test_single_pixel_sample_is_a_hat():
    # linear interpolation of one non-vacuum pixel between vacuum neighbours is
    # the hat max(0, 1 - |x - x_k| / px) times its delta+i*beta; with an even
    # row length the pixel k = cols/2 sits at x = x_position
    p = _sample_params()
    px, pz, db = 6e-8, 1.5e-6, complex(4e-7, 3e-9)
    for xpos in (0.0, 1.3e-7):
        grid = np.zeros((1, 8), dtype=np.int8)
        grid[0, 4] = 1
        s = bw.Sample(z_start=1e-3, pixel_size_x=px, pixel_size_z=pz, grid=grid,
                      db_list=np.array([0, db], dtype=np.complex128), x_positions=np.array([xpos]))
        f = bw.Fields(np.ones(p.N, dtype=np.complex128))
        bw.sample_apply(f, s, p, 1e30, 0)
        x = (np.arange(p.N) - p.N / 2) * p.dx
        hat = np.maximum(0.0, 1.0 - np.abs(x - xpos) / px)
        ref = bw.Fields(np.exp(2j * np.pi * pz / p.wl * db * hat))
        bw.propagate(ref, p.dx, p.wl, pz, p.chunk_size, 1e30)
        assert np.abs(f.u - ref.u).max() < 1e-12


def test_sample_x_position_moves_the_sample_to_plus_x():
    # on a uniform field propagation commutes with shifts, so moving the
    # sample by s samples (x_position = s*dx) rolls the result by +s
    p = _sample_params()
    rng = np.random.default_rng(3)
    grid = rng.integers(0, 3, size=(2, 30)).astype(np.int8)
    grid[:, 0] = grid[:, -1] = 0  # vacuum ends: the modulation is continuous
    dbl = np.array([0, complex(3e-7, 1e-9), complex(8e-7, 6e-9)])
    out = {}
    for sft in (0, 37):
        s = bw.Sample(z_start=1e-3, pixel_size_x=5e-8, pixel_size_z=1e-6, grid=grid, db_list=dbl,
                      x_positions=np.array([sft * p.dx]))
        f = bw.Fields(np.ones(p.N, dtype=np.complex128))
        bw.sample_apply(f, s, p, 1e30, 0)
        out[sft] = f.u
    assert np.abs(out[37] - np.roll(out[0], 37)).max() < 1e-9
    assert np.abs(out[37] - np.roll(out[0], -37)).max() > 1e-3


def test_material_factor_attenuates_and_shifts():
    t, wl = 1e-6, 2e-11
    f = bw.material_factor(complex(1e-6, 1e-8), t, wl)
    assert np.isclose(abs(f), np.exp(-2 * np.pi * t / wl * 1e-8))
    assert np.isclose(np.angle(f), (2 * np.pi * t / wl * 1e-6 + np.pi) % (2 * np.pi) - np.pi)
