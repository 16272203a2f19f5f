"""Checks that pin the oracle's history and keypoints (big-wave/history.py,
wavesim.py:35-58, 153-219).

No reference test or vector covers them, so they are pinned by what the
definitions fix:
  * recording a history does not change the detector rows (each entry is
    propagated from the segment start; the field ends as after one step);
  * free-space entries of a Gaussian beam are the closed-form Fresnel
    Gaussian at each entry's z, binned by square_and_downsample (pinned by
    big-wave's detector KAT), on the z grid arange(0, dz, hdz) with the
    half-step rule for the last entry;
  * the entry after a grating's last slice is the detector row of the same
    setup with the detector right behind the grating, and a sample's first
    entry is the detector row right in front of the sample;
  * a point source's analytic rows px py cos(angle) / r^2 agree with the
    binned intensity of the analytic field 1/r inside the cutoff cone;
  * keypoint (0, 0) of a point source without cutoff is the analytic field.
"""
import math

import numpy as np

from oracle import bigwave as bw


def _params(N, dx, z_det, size, px, wl, chunk=None):
    return bw.SimParams(N=N, dx=dx, z_detector=z_det, detector_size=size, detector_pixel_size_x=px,
                        detector_pixel_size_y=1.0, wl=wl, chunk_size=chunk or N)


def _gauss(N, dx, s, wl, z):
    x = (np.arange(N) - N / 2) * dx
    a = 2 * np.pi ** 2 * s ** 2 - 1j * np.pi * wl * z
    return (np.sqrt(2 * np.pi) * s * np.sqrt(np.pi / a) * np.exp(-np.pi ** 2 * x ** 2 / a)
            * np.exp(-2j * np.pi * z / wl))


def _grating(z0, x_positions=(0.0,), steps=4, sub=0.0):
    return bw.Grating(pitch=1.0e-6, dc=(0.45, 0.55), z_start=z0, thickness=4e-6, nr_steps=steps,
                      x_positions=np.array(x_positions), substrate_thickness=sub,
                      db_a=complex(5.1e-7, 1.6e-10), db_substrate=complex(2e-7, 1e-10))


def test_free_space_history_is_the_gaussian_closed_form():
    N, dx, s, wl = 2 ** 14, 2e-8, 4e-7, 2.7e-11
    p = _params(N, dx, 0.05, N * dx / 2, 4 * dx, wl)
    x = (np.arange(N) - N / 2) * dx
    src = bw.VectorSource(vec=np.exp(-x ** 2 / (2 * s ** 2)).astype(np.complex128), z=0.0)
    h = bw.History(0.011)
    rows = bw.run_simulation(p, src, [], [math.pi / 2], history=h)
    # arange(0, 0.05, 0.011) = 0, .011, .022, .033, .044; 0.05 - 0.044 >= 0.0055
    assert np.allclose(h.zs, [0.011, 0.022, 0.033, 0.044, 0.05], rtol=0, atol=1e-15)
    hm = h.get_history()
    assert hm.shape == (len(bw.detector_x_vector(p.detector_size, p.detector_pixel_size_x)), 5)
    for k, z in enumerate(h.zs):
        ref = bw.square_and_downsample(_gauss(N, dx, s, wl, z), p, z)
        assert np.linalg.norm(hm[:, k] - ref) / np.linalg.norm(ref) < 1e-9
    assert np.array_equal(hm[:, -1], rows[0])


def test_last_step_within_half_a_step_is_moved_to_the_end():
    N, dx, wl = 2 ** 12, 2e-8, 2.7e-11
    p = _params(N, dx, 0.1, N * dx / 2, 4 * dx, wl)
    src = bw.VectorSource(vec=np.ones(N, dtype=np.complex128), z=0.0)
    h = bw.History(0.03)
    bw.run_simulation(p, src, [], [math.pi / 2], history=h)
    # arange(0, 0.1, 0.03) = 0, .03, .06, .09; 0.1 - 0.09 < 0.015 -> last becomes 0.1
    assert np.allclose(h.zs, [0.03, 0.06, 0.1], rtol=0, atol=1e-15)


def test_history_does_not_change_the_detector_rows():
    N, dx, wl = 2 ** 14, 1e-9, 4.1e-11
    p = _params(N, dx, 0.05, 8e-6, 2e-7, wl)
    els = [_grating(0.01, x_positions=(0.0, 2.5e-7), sub=2e-6), _grating(0.02)]
    src = bw.PointSource(x=1e-8, z=0.0)
    ang = [math.pi / 2] * 3
    plain = bw.run_simulation(p, src, els, ang)
    h = bw.History(0.004)
    withh = bw.run_simulation(p, src, els, ang, history=h)
    assert all(np.array_equal(a, b) for a, b in zip(plain, withh))
    # analytic rows at .004, .008; 4 slices; hop .010006 -> .02 in 2 entries
    # (the third step is within half a step); 4 slices; 7 entries to 0.05;
    # nothing from phase step 1
    assert len(h.zs) == 2 + 4 + 2 + 4 + 7


def test_last_slice_entry_is_the_detector_right_behind_the_grating():
    N, dx, wl = 2 ** 14, 1e-9, 4.1e-11
    g = _grating(0.01, steps=5)
    src = bw.PointSource(x=0.0, z=0.0)
    p = _params(N, dx, 0.05, 8e-6, 2e-7, wl)
    h = bw.History(0.004)
    bw.run_simulation(p, src, [g], [math.pi / 2] * 2, history=h)
    z_end = g.z_start + g.thickness
    k = int(np.argmin(np.abs(np.asarray(h.zs) - z_end)))
    behind = bw.run_simulation(_params(N, dx, z_end, 8e-6, 2e-7, wl), src, [g], [math.pi / 2] * 2)
    assert np.array_equal(h.get_history()[:, k], behind[0])
    # slice entries at z_start + (i + 1) * thickness / nr_steps
    zs = np.asarray(h.zs)
    assert np.allclose(zs[k - 4:k + 1], g.z_start + np.arange(1, 6) * g.thickness / 5, rtol=0, atol=1e-18)


def test_sample_first_entry_is_the_detector_in_front_of_the_sample():
    N, dx, wl = 2 ** 14, 1e-9, 4.1e-11
    grid = np.array([[0, 1, 1, 0], [1, 0, 1, 1]], dtype=np.int8)
    smp = bw.Sample(z_start=0.01, pixel_size_x=2e-6, pixel_size_z=3e-6, grid=grid,
                    db_list=np.array([0, complex(3e-7, 1e-9)]), x_positions=np.array([0.0]))
    src = bw.PointSource(x=0.0, z=0.0)
    h = bw.History(0.0035)
    bw.run_simulation(_params(N, dx, 0.03, 8e-6, 2e-7, wl), src, [smp], [math.pi / 2] * 2, history=h)
    # the source reaches the sample analytically: entries .0035, .007, then
    # one before each of the two rows, then the hop to the detector
    assert h.zs[:4] == [0.0035, 0.007, 0.01, 0.01 + 3e-6]
    front = bw.run_simulation(_params(N, dx, 0.01, 8e-6, 2e-7, wl), src, [], [math.pi / 2])
    assert np.array_equal(h.get_history()[:, 2], front[0])


def test_point_source_analytic_rows_match_the_binned_field():
    N, dx, wl = 2 ** 16, 1e-9, 4.1e-11
    size, px = 2e-5, 5e-7
    p = _params(N, dx, 0.05, size, px, wl)
    ang = 0.3e-3
    h = bw.History(0.006)
    src = bw.PointSource(x=2e-7, z=0.0)
    bw.run_simulation(p, src, [_grating(0.02)], [ang, math.pi / 2], history=h)
    fc = math.sin(ang) / wl
    hm = h.get_history()
    xk = bw.detector_x_vector(size, px) - src.x
    assert np.allclose(h.zs[:3], [0.006, 0.012, 0.018], rtol=1e-15, atol=0)
    for k, z in enumerate(h.zs[:3]):
        field = bw.propagate_analytically(N, dx, wl, N, z, src.x, np.complex128)
        ref = bw.square_and_downsample(field, p, z)
        inside = np.abs(xk) < 0.8 * fc * wl * z
        assert inside.sum() > 5
        assert np.allclose(hm[inside, k], ref[inside], rtol=2e-3, atol=0)
        assert np.all(hm[np.abs(xk) > fc * wl * z, k] == 0)


def test_keypoint_before_the_first_element_is_the_point_source():
    N, dx, wl = 2 ** 14, 1e-9, 4.1e-11
    p = _params(N, dx, 0.05, 8e-6, 2e-7, wl)
    g = _grating(0.01)
    kp = {}
    bw.run_simulation(p, bw.PointSource(x=3e-8, z=0.0), [g], [math.pi / 2] * 2, keypoints=kp)
    ref = bw.propagate_analytically(N, dx, wl, N, 0.01, 3e-8, np.complex128)
    assert sorted(kp) == [(0, 0), (0, 1)]
    assert np.linalg.norm(kp[(0, 0)] - ref) / np.linalg.norm(ref) < 1e-12
