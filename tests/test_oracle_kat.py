"""Known-answer tests of the oracle, transcribed from the reference's own suites
(big-wave/test.py and fast-wave/test/fwtest.cu).  These pin the oracle's
primitives independently of its own formulas."""
import numpy as np
import pytest

from oracle import bigwave as bw


def test_fftfreq_chunks_match_numpy():
    # big-wave/test.py:90-97
    for n in [4, 5, 8, 9]:
        expected = np.fft.fftfreq(n)
        for start in range(n):
            for end in range(start, n):
                chunk = bw.fftfreq_chunk(n, start, end - start)
                assert len(chunk) == end - start
                assert np.allclose(expected[start:end], chunk)


def test_fftfreq_fastwave_kat():
    # fast-wave/test/fwtest.cu:59-69 (N=10, dx=1)
    exp = [0.0, 0.1, 0.2, 0.3, 0.4, -0.5, -0.4, -0.3, -0.2, -0.1]
    assert np.allclose(bw.fftfreq_chunk(10, 0, 10), exp, atol=1e-7)


def _grating(pitch, dc, xpos):
    return bw.Grating(pitch=pitch, dc=dc, z_start=0, thickness=1, nr_steps=1,
                      x_positions=np.array(xpos), substrate_thickness=0.0)


def test_grating_templates():
    # big-wave/test.py:121-158
    n, pitch, dx = 16, 4, 0.5
    for offset in [0, 3]:
        g = _grating(pitch, (0.25, 0.5), [offset * dx + 0.01])
        for zf, template in [(0, [1, 0, 0, 0, 0, 0, 0, 1]), (1, [1, 1, 0, 0, 0, 0, 1, 1])]:
            expected = np.tile(template, 10)
            for start in range(n):
                for end in range(start, n):
                    got = bw.grating_mask(0.5, start, end - start, g, zf, 0)
                    assert list(expected[start + offset:end + offset]) == list(got)


def test_grating_stepping_invariance():
    # big-wave/test.py:160-185: shifts by whole pitches leave the mask unchanged
    pitch = 4.2e-6
    g = _grating(pitch, (0.25, 0.5), [0.0, pitch * 0.5, pitch, pitch * 2])
    masks = [bw.grating_mask(1e-6, 123, 16, g, 0, s).tolist() for s in range(4)]
    assert masks[0] == masks[2] == masks[3]
    assert masks[0] != masks[1]


def test_grating_fastwave_kat():
    # fast-wave/test/fwtest.cu:71-87: N=16, dx=1, pitch 6, dc .5, x=(i-8)
    g = _grating(6.0, (0.5, 0.5), [0.0])
    m = bw.grating_mask(1.0, -8, 16, g, 0, 0)
    exp = [0, 1, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0, 0, 1, 1, 1]
    assert list(m.astype(int)) == exp


def test_frequency_cutoff_matches_strict_band():
    # big-wave/test.py:187-199
    for dx in [0.5, 1.0, 2.3]:
        for n in [3, 4, 5, 6]:
            for freq in [0.0, 0.1, 0.4, 1.2, 3.5, 100000.0]:
                v = np.ones(n, dtype=np.complex128)
                bw.apply_frequency_cutoff(v, freq, dx, 3)
                ref = np.ones(n, dtype=np.complex128) * (np.abs(np.fft.fftfreq(n) / dx) < freq)
                assert list(v) == list(ref)


def test_energy_conversion():
    # big-wave/test.py:201-207
    assert abs(bw.convert_energy_wavelength(46000) - 26.95e-12) < 0.01e-12
    assert abs(bw.convert_energy_wavelength(26.95e-12) - 46000) < 10


def test_square_and_downsample_kat():
    # big-wave/test.py:209-225
    p = bw.SimParams(N=8, dx=0.5, z_detector=1.0, detector_size=3.0, detector_pixel_size_x=1.0,
                     detector_pixel_size_y=1.0, wl=1.0, chunk_size=3)
    got = bw.square_and_downsample(np.arange(8, dtype=np.complex128), p, 1.0)
    assert list(got) == [1.25, 12.5, 15.249999999999998]


def test_detector_x_consistency():
    # big-wave/test.py:227-250
    v = np.arange(118, dtype=np.complex128)
    for px in [0.3, 0.5, 0.6, 1.0]:
        for size in [1.0, 2.0, 10.0]:
            p = bw.SimParams(N=118, dx=0.2, z_detector=1.0, detector_size=size,
                             detector_pixel_size_x=px, detector_pixel_size_y=1.0, wl=1.0,
                             chunk_size=3)
            out = bw.square_and_downsample(v, p, 1.0)
            x = bw.detector_x_vector(size, px)
            assert len(out) == len(x)
            assert x[len(x) // 2] == 0
            assert abs(x[0] + x[-1]) <= px * 1.01


def test_stepping_shapes():
    # big-wave/test.py:252-330 (without history): 3 phase steps, 2 pixels
    N = 6
    def grating(z, xp):
        return bw.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=z, thickness=100e-6, nr_steps=1,
                          x_positions=np.array(xp), substrate_thickness=0.0)
    p = bw.SimParams(N=N, dx=10e-9, z_detector=1.0, detector_size=40e-9, detector_pixel_size_x=20e-9,
                     detector_pixel_size_y=1.0, wl=bw.convert_energy_wavelength(46000), chunk_size=3)
    rows = bw.run_simulation(p, bw.PointSource(0.0, 0.0),
                             [grating(0.1, [0.0]), grating(0.2, [0.0, 1.6e-6, 3.1e-6])],
                             [0.1, 0.2, 0.3])
    assert len(rows) == 3 and all(r.shape == (2,) for r in rows)
    rows = bw.run_simulation(p, bw.PointSource(0.0, 0.0),
                             [grating(0.1, [0.0]), grating(0.2, [0.0, 1.6e-6, 3.1e-6]),
                              grating(0.3, [0.0])], [0.1, 0.2, 0.3, 0.4])
    assert len(rows) == 3 and all(r.shape == (2,) for r in rows)


def test_deltabeta_table_lookup():
    # big-wave/test.py:332-351
    table = [(("Au", 3.51), 1), (("Si", 2.34), 2), (("Si", 2.12), 3)]
    assert bw.lookup_deltabeta("Si", 2.34, table) == 2
    assert bw.lookup_deltabeta("Au", 3.51, table) == 1
    with pytest.raises(KeyError):
        bw.lookup_deltabeta("Xy", 1.34, table)


def test_max_dx_monotone():
    # big-wave/test.py:353-366 via grid_density_check: a farther source needs a finer grid
    wl, dz, N = 1.8e-11, 0.1, 2 ** 28

    def max_dx(xs):
        lo, hi = 0.0, 0.1
        while (hi - lo) / hi > 0.01:
            mid = (hi + lo) / 2
            try:
                bw.grid_density_check(dz, xs, N, mid, wl)
                lo = mid
            except ValueError:
                hi = mid
        return lo
    assert max_dx(0.01) < max_dx(0.0)
