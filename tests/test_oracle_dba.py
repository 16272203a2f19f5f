"""Checks that pin the oracle's DBA grating (big-wave/optical_element.py:227-375).

No reference test or vector covers the DBA grating, so the restatement is
pinned by what the definition fixes:
  * with dc1 >= 1 the second condition always holds, and the element is the
    plain grating of (pitch0, dc0) bit for bit (that one is pinned by big-wave's
    grating templates);
  * with dc1 = 0 it never holds, and the element is a uniform slab of material B;
  * the second condition, mod(x / pitch1, 1) < dc1, agrees with exact rational
    arithmetic on the same float x at every sample not within 1e-12 of an edge,
    on each slice of an interpolated duty cycle, and the mask is the AND of the
    two conditions (an OR, a dc/2 offset on the second pitch or a slice index
    off by one fails one of these).
"""
from fractions import Fraction

import numpy as np

from oracle import bigwave as bw

N, DX, WL = 4096, 2e-9, 2.5e-11


def _params(chunk=N):
    return bw.SimParams(N=N, dx=DX, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                        detector_pixel_size_y=1.0, wl=WL, chunk_size=chunk)


def _dba(dc0=(0.5, 0.5), dc1=(0.5, 0.5), pitch0=4.2e-7, pitch1=1.26e-6, steps=4, xpos=(0.0,),
         sub=0.0):
    return bw.DBAGrating(pitch0=pitch0, pitch1=pitch1, dc0=dc0, dc1=dc1, z_start=0.0,
                         thickness=8e-6, nr_steps=steps, x_positions=np.array(xpos),
                         substrate_thickness=sub, db_a=complex(7.7e-7, 3.5e-10),
                         db_b=complex(1e-8, 1e-11), db_substrate=complex(2e-7, 1e-10))


def _field():
    rng = np.random.default_rng(5)
    return rng.standard_normal(N) + 1j * rng.standard_normal(N)


def _exact_second(x: np.ndarray, pitch1: float, dc1: float):
    """(decision, near_edge) of mod(x/pitch1, 1) < dc1 in exact arithmetic."""
    p, d = Fraction(pitch1), Fraction(dc1)
    dec = np.zeros(len(x), dtype=bool)
    near = np.zeros(len(x), dtype=bool)
    eps = Fraction(1, 10 ** 12)
    for i, xv in enumerate(x):
        q = Fraction(float(xv)) / p
        fr = q - (q.numerator // q.denominator)
        dec[i] = fr < d
        near[i] = fr < eps or abs(fr - d) < eps or 1 - fr < eps
    return dec, near


def test_dba_with_full_second_duty_cycle_is_the_plain_grating():
    p = _params(chunk=1000)
    for xpos in (0.0, 1.05e-7):
        d = _dba(dc0=(0.3, 0.6), dc1=(1.5, 1.5), xpos=(xpos,), sub=3e-6)
        g = bw.Grating(pitch=d.pitch0, dc=d.dc0, z_start=0.0, thickness=d.thickness,
                       nr_steps=d.nr_steps, x_positions=d.x_positions,
                       substrate_thickness=d.substrate_thickness, db_a=d.db_a, db_b=d.db_b,
                       db_substrate=d.db_substrate)
        u = _field()
        fd, fg = bw.Fields(u.copy()), bw.Fields(u.copy())
        bw.dba_grating_apply(fd, d, p, 1e30, 0)
        bw.grating_apply(fg, g, p, 1e30, 0)
        assert np.array_equal(fd.u, fg.u)


def test_dba_with_zero_second_duty_cycle_is_a_slab_of_b():
    p = _params()
    d = _dba(dc1=(0.0, 0.0))
    g = bw.Grating(pitch=d.pitch0, dc=d.dc0, z_start=0.0, thickness=d.thickness,
                   nr_steps=d.nr_steps, x_positions=d.x_positions, substrate_thickness=0.0,
                   db_a=d.db_b, db_b=d.db_b)
    u = _field()
    fd, fg = bw.Fields(u.copy()), bw.Fields(u.copy())
    bw.dba_grating_apply(fd, d, p, 1e30, 0)
    bw.grating_apply(fg, g, p, 1e30, 0)
    assert np.array_equal(fd.u, fg.u)


def test_dba_second_condition_matches_exact_arithmetic_per_slice():
    steps, lo, hi = 4, 0.2, 0.6
    d = _dba(dc0=(1.5, 1.5), dc1=(lo, hi), steps=steps, xpos=(3.3e-8,))
    x = np.linspace(-N // 2 * DX + 3.3e-8, -N // 2 * DX + 3.3e-8 + N * DX, N, endpoint=False)
    checked = 0
    for i in range(steps):
        dc1 = lo + (hi - lo) * i / steps
        mask = bw.dba_grating_mask(DX, -N // 2, N, d, i / steps, 0)
        dec, near = _exact_second(x, d.pitch1, dc1)
        assert np.array_equal(mask[~near], dec[~near])
        checked += int((~near).sum())
        # the duty cycle of slice i, not of slice i +- 1
        assert abs(mask.mean() - dc1) < 0.05
    assert checked > steps * N * 0.99


def test_dba_mask_is_the_and_of_both_conditions():
    d = _dba(dc0=(0.4, 0.7), dc1=(0.35, 0.55), steps=3, xpos=(-2.1e-7,))
    g = bw.Grating(pitch=d.pitch0, dc=d.dc0, z_start=0.0, thickness=d.thickness,
                   nr_steps=d.nr_steps, x_positions=d.x_positions, substrate_thickness=0.0)
    x = np.linspace(-N // 2 * DX - 2.1e-7, -N // 2 * DX - 2.1e-7 + N * DX, N, endpoint=False)
    for i in range(3):
        zf = i / 3
        first = bw.grating_mask(DX, -N // 2, N, g, zf, 0)
        dc1 = (1 - zf) * d.dc1[0] + zf * d.dc1[1]
        second, near = _exact_second(x, d.pitch1, dc1)
        mask = bw.dba_grating_mask(DX, -N // 2, N, d, zf, 0)
        assert np.array_equal(mask[~near], (first & second)[~near])
        # both conditions matter on this geometry
        assert (first & ~second).any() and (second & ~first).any()
