"""GPU parity of the HIP path against the complex128 oracle (big-wave
restatement) on the same seeded inputs.

Tolerances: complex128 fields (c16) to ~1e-12 relative; complex64 fields
with FP64 phase arguments (c8) to the 1e-4 relative-L2 intensity bound and
1e-3 rad fringe phase of the precision form; integer / mask / pixel-range
work bit exact."""
import math

import numpy as np
import pytest

from oracle import bigwave as bw
from simcases import DB_SI_25KEV, mono_chain, talbot_lau, sample_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import ravesim
    c = ravesim.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rand_field(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(N) + 1j * rng.standard_normal(N)


# N covering one, two and three digit groups for each precision
PROP_CASES = [(2 ** 10, np.complex128), (2 ** 12, np.complex64), (2 ** 16, np.complex128),
              (2 ** 18, np.complex64), (2 ** 22, np.complex128), (2 ** 23, np.complex64)]


@pytest.mark.parametrize("N,dtype", PROP_CASES)
@pytest.mark.parametrize("plain", [False, True])
def test_propagate_matches_oracle(ctx, N, dtype, plain):
    import ravesim
    dx, wl, dz = 2e-9, 2.7e-11, 1.7e-3
    u = rand_field(N, N)
    fc = 0.3 / dx  # removes part of the band
    f = ravesim.Field(ctx, N, dx, dtype=dtype, plain=plain)
    f.upload(u)
    f.propagate(wl, dz, fc)
    got = f.download()
    ref = bw.Fields(u.astype(np.complex128))
    bw.propagate(ref, dx, wl, dz, N, fc)
    tol = 1e-10 if dtype == np.complex128 else 2e-6
    assert rel_l2(got, ref.u) < tol


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_two_hops_equal_one(ctx, dtype):
    import ravesim
    N, dx, wl = 2 ** 20, 2e-9, 2.7e-11
    u = rand_field(N, 5)
    a = ravesim.Field(ctx, N, dx, dtype=dtype)
    a.upload(u)
    a.propagate(wl, 0.03)
    a.propagate(wl, 0.07)
    b = ravesim.Field(ctx, N, dx, dtype=dtype)
    b.upload(u)
    b.propagate(wl, 0.10)
    tol = 1e-10 if dtype == np.complex128 else 1e-5
    assert rel_l2(a.download(), b.download()) < tol


@pytest.mark.parametrize("N,chunk,xpos", [(2 ** 14, None, 0.0), (2 ** 16, 1000, 3.3e-7),
                                          (2 ** 20, 10469376 // 16, -1.05e-6)])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_grating_mask_bit_exact(ctx, N, chunk, xpos, dtype):
    """pitch/dx = 2100 puts samples exactly on mask edges; multi-chunk x."""
    import ravesim
    dx, wl = 2e-9, 2.7e-11
    t = 1e-12  # negligible propagation; material A halves the amplitude
    beta = math.log(2) * wl / (2 * math.pi * t)
    g = ravesim.Grating(pitch=4.2e-6, dc=(0.37, 0.62), z_start=0.0, thickness=t * 3, nr_steps=3,
                        x_positions=[xpos, xpos + 1e-7], db_a=complex(0, beta))
    chunk = chunk or N
    f = ravesim.Field(ctx, N, dx, dtype=dtype, chunk_size=chunk)
    f.upload(np.ones(N))
    og = bw.Grating(pitch=4.2e-6, dc=(0.37, 0.62), z_start=0.0, thickness=t * 3, nr_steps=3,
                    x_positions=np.array([xpos, xpos + 1e-7]), substrate_thickness=0.0)
    # one slice at a time: compare the mask of each slice separately
    g1 = ravesim.Grating(pitch=4.2e-6, dc=(0.37, 0.37), z_start=0.0, thickness=t, nr_steps=1,
                         x_positions=[xpos, xpos + 1e-7], db_a=complex(0, beta))
    f.apply_grating(wl, g1, phase_step=1)
    got = np.abs(f.download()) < 0.75
    exp = np.concatenate([bw.grating_mask(dx, i - N // 2, min(chunk, N - i), og, 0.0, 1)
                          for i in range(0, N, chunk)])
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("N,xpos,dtype", [(2 ** 14, 0.0, np.complex64), (2 ** 20, -1.05e-6, np.complex128),
                                           (2 ** 26, 3.3e-7, np.complex64), (2 ** 25, 0.0, np.complex128)])
def test_grating_mask_cached_bit_exact(ctx, N, xpos, dtype):
    """Slices of a constant-duty-cycle grating share one mask.  Fields larger
    than the 256 MiB last-level cache (the last two cases) get it computed once
    per run and read back as bits in every slice's top pass; smaller ones
    decide inline.  Both stay big-wave's decisions at exact ties (pitch/dx =
    2100)."""
    import ravesim
    dx, wl = 2e-9, 2.7e-11
    t = 1e-12
    beta = math.log(2) * wl / (2 * math.pi * t)  # each slice halves the amplitude in A
    g = ravesim.Grating(pitch=4.2e-6, dc=(0.41, 0.41), z_start=0.0, thickness=4 * t, nr_steps=4,
                        x_positions=[xpos], db_a=complex(0, beta))
    f = ravesim.Field(ctx, N, dx, dtype=dtype)
    f.upload(np.ones(N))
    f.apply_grating(wl, g, phase_step=0)
    a = np.abs(f.download())
    f.close()
    og = bw.Grating(pitch=4.2e-6, dc=(0.41, 0.41), z_start=0.0, thickness=t, nr_steps=1,
                    x_positions=np.array([xpos]), substrate_thickness=0.0)
    exp = bw.grating_mask(dx, -(N // 2), N, og, 0.0, 0)
    got = a < 0.25  # 1/16 inside A, ~1 outside
    assert np.array_equal(got, exp)
    assert np.all(np.abs(a[exp] - 1.0 / 16.0) < 1e-3) and np.all(np.abs(a[~exp] - 1.0) < 1e-3)


@pytest.mark.parametrize("N,dtype", [(64, np.complex64), (2 ** 20, np.complex128), (2 ** 26, np.complex64),
                                     (2 ** 25, np.complex128)])
def test_upload_download_round_trip(ctx, N, dtype):
    """rs_field_upload / rs_field_download go through the context's pinned
    staging buffers (one to four copying threads by size): the field comes
    back bit for bit, and a copy on the device matches too."""
    import ravesim
    rng = np.random.default_rng(N)
    u = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(dtype)
    f = ravesim.Field(ctx, N, 2e-9, dtype=dtype)
    f.upload(u)
    assert np.array_equal(f.download(), u)
    g = ravesim.Field(ctx, N, 2e-9, dtype=dtype)
    f.copy_to(g)
    f.upload(np.zeros(N, dtype=dtype))
    assert np.array_equal(g.download(), u)
    assert not np.any(f.download())
    f.close()
    g.close()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_point_source_matches_oracle(ctx, dtype):
    import ravesim
    N, dx, wl = 2 ** 16, 3.1e-10, 2.67e-11
    f = ravesim.Field(ctx, N, dx, dtype=dtype)
    f.point_source(wl, 4.4e-7, 0.1)
    ref = bw.propagate_analytically(N, dx, wl, N, 0.1, 4.4e-7)
    tol = 1e-13 if dtype == np.complex128 else 1e-6
    assert rel_l2(f.download(), ref) < tol


def test_detector_known_answer(ctx):
    """big-wave/test.py:209-225 on an 8-sample field; fields here need
    N >= 64, so the 8 samples sit at the same x positions (n = -4..3) of a
    64-sample field whose other samples are zero."""
    import ravesim
    u = np.zeros(64, dtype=np.complex128)
    u[28:36] = np.arange(8)
    f = ravesim.Field(ctx, 64, 0.5, dtype=np.complex128)
    f.upload(u)
    got = f.square_and_downsample(3.0, 1.0, 1.0, 1.0)
    assert np.allclose(got, [1.25, 12.5, 15.249999999999998], rtol=1e-14, atol=0)


@pytest.mark.parametrize("N,size,px", [(2 ** 16, 1e-4, 1e-6), (2 ** 16, 5e-5, 3e-7), (2 ** 12, 4e-6, 1e-7)])
def test_detector_matches_histogram(ctx, N, size, px):
    import ravesim
    dx = 2e-9
    u = rand_field(N, 11)
    f = ravesim.Field(ctx, N, dx, dtype=np.complex128)
    f.upload(u)
    got = f.square_and_downsample(size, px, 1.0, 1.0)
    p = bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=size, detector_pixel_size_x=px,
                     detector_pixel_size_y=1.0, wl=1.0, chunk_size=N)
    exp = bw.square_and_downsample(u, p, 1.0)
    assert got.shape == exp.shape
    assert np.allclose(got, exp, rtol=1e-12, atol=0)


def _simdir(tmp_path, case):
    from ravesim.simdir import write_simdir
    cfg, subs, ang = case
    return write_simdir(tmp_path / "sim", cfg, subs, ang)


@pytest.mark.parametrize("dtype", ["c16", "c8"])
def test_mono_chain_end_to_end(ctx, tmp_path, dtype):
    from ravesim.multisim import run_single_simulation
    sd = _simdir(tmp_path, mono_chain(tmp_path, N=2 ** 16, slices=8, det_size=1e-4, dtype=dtype))
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    assert got.shape == ref.shape == (1, 100)
    saved = np.load(sd / "00000000" / "detected.npy")
    assert saved.dtype == np.float64 and np.array_equal(saved, got)
    tol = 1e-11 if dtype == "c16" else 1e-4
    assert rel_l2(got, ref) < tol


@pytest.mark.parametrize("N,file_dtype,dtype", [(2 ** 16, np.complex64, "c8"), (2 ** 16, np.complex128, "c8"),
                                                (2 ** 23, np.complex64, "c16"), (2 ** 21, np.complex128, "c16"),
                                                (2 ** 24, np.complex64, "c8")])
def test_vector_source_streams_from_npy(ctx, tmp_path, N, file_dtype, dtype):
    """A vector source is read from its .npy in chunks through pinned staging
    buffers, by one host thread per 32 MiB of field up to four (several chunks
    per thread and both of its buffers at 2^21 and above, with conversion
    between complex64 files and complex128 fields and back): the detector rows
    match the oracle, which loads the whole file; a truncated file is an
    error before anything runs."""
    from ravesim import _lib
    from ravesim.multisim import run_single_simulation
    cfg, subs, ang = mono_chain(tmp_path, N=N, slices=2, det_size=min(1e-4, N * 2e-9 / 2), dtype=dtype)
    rng = np.random.default_rng(N)
    u = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(file_dtype)
    src = tmp_path / "src.npy"
    np.save(src, u)
    subs[0]["source"] = {"type": "vector", "input_path": str(src), "z": 0.0}
    sd = _simdir(tmp_path, (cfg, subs, ang))
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    tol = 1e-10 if dtype == "c16" else 1e-4
    assert rel_l2(got, ref) < tol
    raw = src.read_bytes()
    src.write_bytes(raw[:-8])
    with pytest.raises(_lib.RsError, match="truncated"):
        run_single_simulation(sd, 0, tmp_path, ctx=ctx)


def test_talbot_lau_phase_stepping(ctx, tmp_path):
    from ravesim.multisim import run_single_simulation
    sd = _simdir(tmp_path, talbot_lau(tmp_path, N=2 ** 18, slices=6, det_size=1e-4, dtype="c8", g2_steps=8))
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    assert got.shape == ref.shape == (8, 100)
    assert rel_l2(got, ref) < 1e-4
    k = np.arange(8)
    def phase(I):
        a_s = (I * np.sin(2 * np.pi * k / 8)[:, None]).sum(0)
        a_c = (I * np.cos(2 * np.pi * k / 8)[:, None]).sum(0)
        return -np.arctan2(a_s, a_c)
    dphi = np.angle(np.exp(1j * (phase(got) - phase(ref))))
    assert np.abs(dphi).max() < 1e-3


def test_plain_and_fused_agree(ctx, tmp_path):
    from ravesim.multisim import run_single_simulation
    sd = _simdir(tmp_path, talbot_lau(tmp_path, N=2 ** 17, slices=3, det_size=1e-4, dtype="c16", g2_steps=2))
    a = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    b = run_single_simulation(sd, 0, tmp_path, ctx=ctx, plain=True)
    assert rel_l2(a, b) < 1e-11


def test_talbot_lau_194_step_error_budget(ctx, tmp_path):
    """All 194 propagation steps of the Talbot-Lau chain (3 x 64 slices) at
    reduced N: the complex64 + FP64-phase form stays within 1e-4 relative L2
    of detector intensity against the complex128 oracle."""
    from ravesim.multisim import run_single_simulation
    sd = _simdir(tmp_path, talbot_lau(tmp_path, N=2 ** 18, slices=64, det_size=1e-4, dtype="c8"))
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    err = rel_l2(got, ref)
    print(f"Talbot-Lau N=2^18 194 steps c8 rel-L2 = {err:.3e}")
    assert err < 1e-4


# digit-group splits forced through rs_set_split: wide column passes
# (C * L = 16384 complex64 / 8192 complex128 per workgroup, XCD-grouped
# partial-line segments) and three-group splits with uneven columns
SPLIT_CASES = [(2 ** 20, np.complex64, 10, 10), (2 ** 22, np.complex64, 9, 13),
               (2 ** 24, np.complex64, 12, 12), (2 ** 20, np.complex128, 8, 12),
               (2 ** 21, np.complex128, 10, 11), (2 ** 22, np.complex64, 8, 5)]


@pytest.mark.parametrize("N,dtype,lgR,lgA", SPLIT_CASES)
def test_forced_splits_match_oracle(N, dtype, lgR, lgA):
    import ravesim
    c = ravesim.Context(0)
    c.set_split(N, dtype, lgR, lgA)
    dx, wl, dz = 2e-9, 2.7e-11, 1.7e-3
    u = rand_field(N, lgR * 100 + lgA)
    fc = 0.3 / dx
    f = ravesim.Field(c, N, dx, dtype=dtype)
    f.upload(u)
    f.propagate(wl, dz, fc)
    got = f.download()
    ref = bw.Fields(u.astype(np.complex128))
    bw.propagate(ref, dx, wl, dz, N, fc)
    tol = 1e-10 if dtype == np.complex128 else 2e-6
    assert rel_l2(got, ref.u) < tol
    f.close()
    c.close()


def test_pass_timing_samples_every_kth_launch():
    """rs_set_timing(k) times launches 0, k, 2k, ... of each pass class; the
    algorithmic bytes of a timed launch are one read and one write of the
    samples it covers (half the field per middle or row launch over two
    spans); timing does not change results."""
    import ravesim
    from ravesim.core import BatchMember, run_batch
    N, dx = 2 ** 18, 2e-9
    c = ravesim.Context(0)
    c.set_split(N, np.complex64, 8, 6)
    g = ravesim.Grating(pitch=1.3e-6, dc=(0.45, 0.45), z_start=0.0, thickness=8e-6, nr_steps=5,
                        db_a=complex(2e-6, 3e-8))
    ms = [BatchMember(wl=ravesim.energy_to_wavelength(30000.0), gratings=[g], cutoff_angles=[0.2] * 2,
                      plane_wave=True)]
    kw = dict(det_size=2e-4, det_px=1e-6, det_py=1.0, z_detector=0.05, dtype=np.complex64, spans=2)
    base = run_batch(c, N, dx, ms, **kw)
    counts = {}
    for k in (1, 3):
        c.set_timing(True, stride=k)
        assert np.array_equal(run_batch(c, N, dx, ms, **kw), base)
        st = c.pass_stats()
        c.set_timing(False)
        counts[k] = {cls: st[cls][0] for cls in ("row", "top", "mid")}
        for cls in ("row", "top", "mid"):
            n, ms_, byts = st[cls]
            assert n > 0 and ms_ > 0
            per = N * 8 * 2 // (1 if cls == "top" else 2)
            assert byts == n * per, cls
    for cls in ("row", "top", "mid"):
        assert counts[3][cls] == -(-counts[1][cls] // 3), cls
    with pytest.raises(ValueError):
        c.set_timing(True, stride=0)
    c.close()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_graph_capture_matches_eager(dtype):
    """rs_batch_desc.use_graph captures the whole run as one hipGraph and
    launches it: three graph runs, with eager runs and synchronisations
    between them, each equal to the eager result (the same kernels on the
    same inputs); one lane (explicit group size) and two lanes (default
    grouping, forked and joined by events inside the capture)."""
    import ravesim
    from ravesim.core import BatchMember, run_batch
    N, dx = 2 ** 16, 2e-9
    c = ravesim.Context(0)
    g = ravesim.Grating(pitch=0.9e-6, dc=(0.45, 0.5), z_start=0.0, thickness=6e-6, nr_steps=3,
                        x_positions=[2e-7], db_a=complex(2e-6, 3e-8))
    ms = [BatchMember(wl=ravesim.energy_to_wavelength(e), gratings=[g], cutoff_angles=[0.2] * 2,
                      source_x=1e-7 * k, source_z=-0.1, weight=0.5 + 0.1 * k)
          for k, e in enumerate((25e3, 31e3, 37e3))]
    for max_resident in (1, 0):
        kw = dict(det_size=5e-5, det_px=1e-6, det_py=1.0, z_detector=0.03, dtype=dtype, max_resident=max_resident)
        eager = run_batch(c, N, dx, ms, **kw)
        for _ in range(3):
            got = run_batch(c, N, dx, ms, use_graph=True, **kw)
            c.synchronize()
            assert np.array_equal(got, eager)
            assert np.array_equal(run_batch(c, N, dx, ms, **kw), eager)
    c.close()


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_streaming_groups_give_identical_rows(dtype):
    """Three 2^24 members run one at a time (a group inside the last-level
    cache: inline grating test, ordinary loads) and as one group of three
    (beyond the cache: cached grating bits and the non-temporal top pass):
    the same samples go through the same arithmetic, and the detector sums
    run in the same member order, so the rows are identical."""
    import ravesim
    from ravesim.core import BatchMember, run_batch
    N, dx = 2 ** 24, 2e-9
    c = ravesim.Context(0)
    c.set_split(N, dtype, 12, 6)  # row 4096 | top 64 | middle 64
    g = ravesim.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=20e-6, nr_steps=4,
                        db_a=complex(7.7e-7, 3.5e-10))
    ms = [BatchMember(wl=ravesim.energy_to_wavelength(e), gratings=[g], cutoff_angles=[0.3] * 2, weight=0.5 + k,
                      plane_wave=True) for k, e in enumerate((30e3, 40e3, 50e3))]
    kw = dict(det_size=5e-4, det_px=1e-6, det_py=1.0, z_detector=0.05, dtype=dtype)
    one = run_batch(c, N, dx, ms, max_resident=1, **kw)
    three = run_batch(c, N, dx, ms, max_resident=3, **kw)
    c.close()
    assert np.array_equal(one, three)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_spans_give_identical_rows(dtype):
    """The middle, row and middle passes of each propagation run over 1..64
    consecutive sub-ranges of the fields (a 2^26 complex64 field runs over two
    by default): every sample goes through the same tile either way, so the
    detector rows are identical, and they match the oracle."""
    import ravesim
    from ravesim import _lib
    from ravesim.core import BatchMember, run_batch
    N, dx = 2 ** 18, 2e-9
    c = ravesim.Context(0)
    c.set_split(N, dtype, 8, 6)  # row 256 | top 64 | middle 16
    wl = ravesim.energy_to_wavelength(30000.0)
    gcfg = [dict(pitch=1.3e-6, dc=(0.45, 0.45), z_start=0.0, thickness=8e-6, nr_steps=3, db_a=complex(2e-6, 3e-8)),
            dict(pitch=1.3e-6, dc=(0.5, 0.5), z_start=0.05, thickness=8e-6, nr_steps=2, db_a=complex(2e-6, 3e-8))]
    gr = [ravesim.Grating(**g) for g in gcfg]
    # lockstep members share one operation structure: two point sources
    ms = [BatchMember(wl=wl, gratings=gr, cutoff_angles=[0.2] * 3, source_x=-1e-7, source_z=-0.2, weight=0.7),
          BatchMember(wl=wl * 1.1, gratings=gr, cutoff_angles=[0.2] * 3, source_x=2e-7, source_z=-0.1)]
    kw = dict(det_size=2e-4, det_px=1e-6, det_py=1.0, z_detector=0.1 + 8e-6, dtype=dtype, max_resident=2)
    base = run_batch(c, N, dx, ms, spans=1, **kw)
    for sp in (2, 8, 64):
        assert np.array_equal(run_batch(c, N, dx, ms, spans=sp, **kw), base), sp
    for bad in (3, 128):
        with pytest.raises(_lib.RsError, match="spans"):
            run_batch(c, N, dx, ms, spans=bad, **kw)
    c.close()
    # the same setup through the oracle, member by member (complex128)
    ogs = [bw.Grating(x_positions=np.array([0.0]), substrate_thickness=0.0, **g) for g in gcfg]
    ref = 0
    for m in ms:
        p = bw.SimParams(N=N, dx=dx, z_detector=kw["z_detector"], detector_size=kw["det_size"],
                         detector_pixel_size_x=kw["det_px"], detector_pixel_size_y=1.0, wl=m.wl, chunk_size=N)
        src = bw.PointSource(x=m.source_x, z=m.source_z)
        ref = ref + m.weight * bw.run_simulation(p, src, ogs, m.cutoff_angles)[0]
    tol = 1e-9 if dtype == np.complex128 else 1e-4
    assert rel_l2(base[0], ref) < tol


def test_autotune_per_batch_size():
    """Splits tuned for one field and for a lockstep batch are both valid
    plans; a batch run under the batch-tuned split matches the default
    split's result (complex64 rounding only)."""
    import ravesim
    from ravesim.core import BatchMember, run_batch
    N, dx = 2 ** 20, 2e-9
    c = ravesim.Context(0)
    one = c.autotune(N, np.complex64, members=1)
    many = c.autotune(N, np.complex64, members=4)
    for t in (one, many):
        assert t["groups"] in (1, 2, 3) and t["ms_per_slice"] > 0
        assert t["lgA"] + t["lgB"] + t["lgR"] == 20 or (t["groups"] == 1 and t["lgR"] == 20)
    ms = []
    for k in range(4):
        g = ravesim.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=30e-6, nr_steps=4,
                            db_a=complex(7.7e-7 / (1 + k), 3.5e-10))
        ms.append(BatchMember(wl=ravesim.energy_to_wavelength((30 + 5 * k) * 1e3), gratings=[g],
                              cutoff_angles=[math.pi / 2] * 2, weight=1.0 + k, plane_wave=True))
    kw = dict(det_size=1e-3, det_px=1e-6, det_py=1.0, z_detector=0.05)
    tuned = run_batch(c, N, dx, ms, **kw)
    c.close()
    d = ravesim.Context(0)
    default = run_batch(d, N, dx, ms, **kw)
    d.close()
    assert rel_l2(tuned, default) < 1e-5


# every row and column fiber length of both precisions through a forced
# two-group split (row 2^lgR, column 2^lgA), N <= 2^21: each Stockham stage
# shape (remainder radix 2/4/8/16, small-stage twiddle tables, powers from
# W_L) runs once against the oracle
FIBER_CASES = ([(np.complex64, r, a) for r in range(6, 14) for a in (4, 7)] +
               [(np.complex64, 6, a) for a in range(10, 14)] +
               [(np.complex128, r, a) for r in range(6, 13) for a in (5, 8)] +
               [(np.complex128, 6, a) for a in range(10, 13)])


@pytest.mark.parametrize("dtype,lgR,lgA", FIBER_CASES)
def test_every_fiber_length(ctx, dtype, lgR, lgA):
    N = 2 ** (lgR + lgA)
    ctx.set_split(N, dtype, lgR, lgA)
    dx, wl, dz = 2e-9, 2.7e-11, 2.3e-3
    u = rand_field(N, 1000 + lgR * 16 + lgA)
    fc = 0.35 / dx
    import ravesim
    f = ravesim.Field(ctx, N, dx, dtype=dtype)
    f.upload(u)
    f.propagate(wl, dz, fc)
    got = f.download()
    f.close()
    ref = bw.Fields(u.astype(np.complex128))
    bw.propagate(ref, dx, wl, dz, N, fc)
    tol = 1e-10 if dtype == np.complex128 else 2e-6
    assert rel_l2(got, ref.u) < tol


# every column length from 16 to 512 as the top group (grating mask inside
# the pass) and as the middle group of a three-group plan, in both
# precisions: each column tile width the passes use
COL_CASES = ([(dt, 6, a, 0) for dt in (np.complex64, np.complex128) for a in range(4, 10)] +
             [(dt, 6, 4, b) for dt in (np.complex64, np.complex128) for b in range(4, 10)])


@pytest.mark.parametrize("dtype,lgR,lgA,lgB", COL_CASES)
def test_every_column_tile(ctx, dtype, lgR, lgA, lgB):
    import ravesim
    N, dx = 2 ** (lgR + lgA + lgB), 2e-9
    ctx.set_split(N, dtype, lgR, lgA)
    wl = ravesim.energy_to_wavelength(30000.0)
    gk = dict(pitch=0.37e-6, dc=(0.4, 0.55), z_start=0.0, thickness=6e-6, nr_steps=3, db_a=complex(2e-6, 4e-8))
    f = ravesim.Field(ctx, N, dx, dtype=dtype)
    f.upload(rand_field(N, 7 * lgA + lgB))
    fc = 0.4 / dx
    f.apply_grating(wl, ravesim.Grating(x_positions=[3.1e-8], **gk), cutoff_freq=fc)
    f.propagate(wl, 1.3e-3, fc)
    got = f.download()
    f.close()
    p = bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                     detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
    ref = bw.Fields(rand_field(N, 7 * lgA + lgB))
    bw.grating_apply(ref, bw.Grating(x_positions=np.array([3.1e-8]), substrate_thickness=0.0, **gk), p, fc, 0)
    bw.propagate(ref, dx, wl, 1.3e-3, N, fc)
    tol = 1e-10 if dtype == np.complex128 else 2e-6
    assert rel_l2(got, ref.u) < tol


def set_plan_tiles(c, path, N, dtype, lgR, lgA, tileA, tileB):
    """Fix split and column tiles for every batch size through a loaded tuning table."""
    c.set_split(N, dtype, lgR, lgA)
    c.tune_save(path)
    dev = next(l for l in path.read_text().splitlines() if not l.startswith("#")).split()[0]
    dt = "c8" if np.dtype(dtype) == np.complex64 else "c16"
    path.write_text(f"{dev} {int(math.log2(N))} {dt} 0 {lgR} {lgA} {tileA} {tileB} 0\n")
    assert c.tune_load(path) == 1
    t = c.tuned(N, dtype, 1)
    assert (t["lgR"], t["lgA"], t["tileA"], t["tileB"]) == (lgR, lgA, tileA, tileB)


@pytest.mark.parametrize("dtype,lgR,lgA,lgB", COL_CASES)
def test_every_column_tile_alternative(ctx, tmp_path, dtype, lgR, lgA, lgB):
    """The autotuner's alternative tile width of every column length, top
    (grating mask inside) and middle, against the oracle."""
    import ravesim
    N, dx = 2 ** (lgR + lgA + lgB), 2e-9
    set_plan_tiles(ctx, tmp_path / "tiles.txt", N, dtype, lgR, lgA, 1 if lgB == 0 else 0, 1 if lgB else 0)
    wl = ravesim.energy_to_wavelength(30000.0)
    gk = dict(pitch=0.37e-6, dc=(0.4, 0.55), z_start=0.0, thickness=6e-6, nr_steps=3, db_a=complex(2e-6, 4e-8))
    u = rand_field(N, 11 * lgA + lgB)
    fc = 0.4 / dx
    f = ravesim.Field(ctx, N, dx, dtype=dtype)
    f.upload(u)
    f.apply_grating(wl, ravesim.Grating(x_positions=[3.1e-8], **gk), cutoff_freq=fc)
    f.propagate(wl, 1.3e-3, fc)
    got = f.download()
    f.close()
    p = bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                     detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
    ref = bw.Fields(u)
    bw.grating_apply(ref, bw.Grating(x_positions=np.array([3.1e-8]), substrate_thickness=0.0, **gk), p, fc, 0)
    bw.propagate(ref, dx, wl, 1.3e-3, N, fc)
    tol = 1e-10 if dtype == np.complex128 else 2e-6
    assert rel_l2(got, ref.u) < tol


def test_tuning_table_persists_across_contexts(tmp_path):
    """A table saved by one context and loaded by another gives the same
    plan (split and tiles) and bit-identical rows; a malformed table is a
    configuration error and entries of another device are skipped."""
    import ravesim
    from ravesim._lib import RS_ERR_CONFIG, RsError
    from ravesim.core import BatchMember, run_batch
    N, dx = 2 ** 18, 2e-9
    ms = [BatchMember(wl=ravesim.energy_to_wavelength((30 + 7 * k) * 1e3),
                      gratings=[ravesim.Grating(pitch=4.2e-6, dc=(0.5, 0.5), z_start=0.0, thickness=30e-6,
                                                nr_steps=3, db_a=complex(7.7e-7, 3.5e-10))],
                      cutoff_angles=[math.pi / 2] * 2, plane_wave=True) for k in range(4)]
    kw = dict(det_size=2e-4, det_px=1e-6, det_py=1.0, z_detector=0.05)
    a = ravesim.Context(0)
    ta = a.autotune(N, np.complex64, members=4)
    path = tmp_path / "tuned.txt"
    a.tune_save(path)
    rows_a = run_batch(a, N, dx, ms, **kw)
    b = ravesim.Context(0)
    assert b.tuned(N, np.complex64, 4) is None
    assert b.tune_load(path) >= 1
    tb = b.tuned(N, np.complex64, 4)
    assert (tb["lgR"], tb["lgA"], tb["tileA"], tb["tileB"]) == (ta["lgR"], ta["lgA"], ta["tileA"], ta["tileB"])
    assert np.array_equal(run_batch(b, N, dx, ms, **kw), rows_a)
    other = tmp_path / "other.txt"
    other.write_text("gfx000:other/1 18 c8 4 9 9 0 0 0.1\n")
    assert b.tune_load(other) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text(path.read_text().splitlines()[1].rsplit(" ", 3)[0] + "\n")
    with pytest.raises(RsError) as e:
        b.tune_load(bad)
    assert e.value.code == RS_ERR_CONFIG
    a.close()
    b.close()


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-10), ("c8", 1e-4)])
def test_sample_element_matches_oracle(ctx, tmp_path, dtype, tol):
    """Sample grid rows (np.interp of delta + i beta between pixel centres,
    zero outside the grid) with phase stepping, against the oracle."""
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    cfg, subs, ang = sample_case(tmp_path, dtype=dtype)
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    import shutil
    shutil.copy(tmp_path / "sample_grid.npy", sd / "sample_grid.npy")
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    assert got.shape == ref.shape == (2, 40)
    err = rel_l2(got, ref)
    print(f"sample element {dtype} rel-L2 = {err:.3e}")
    assert err < tol
    assert rel_l2(got[0], got[1]) > 1e-3  # the two sample positions differ


@pytest.mark.parametrize("N,chunk,xpos,dtype,plain", [
    (2 ** 14, None, 0.0, np.complex64, False), (2 ** 16, 1000, 3.3e-7, np.complex128, False),
    (2 ** 20, None, -1.05e-6, np.complex64, False), (2 ** 16, None, 2.1e-7, np.complex128, True)])
def test_dba_grating_mask_bit_exact(ctx, N, chunk, xpos, dtype, plain):
    """DBA grating decisions, one slice at a time: pitch0/dx = 2100 and
    pitch1/dx = 6300 put samples exactly on both conditions' edges."""
    import ravesim
    dx, wl = 2e-9, 2.7e-11
    t = 1e-12  # negligible propagation; material A halves the amplitude
    beta = math.log(2) * wl / (2 * math.pi * t)
    chunk = chunk or N
    og = bw.DBAGrating(pitch0=4.2e-6, pitch1=12.6e-6, dc0=(0.37, 0.62), dc1=(0.55, 0.3),
                       z_start=0.0, thickness=t * 3, nr_steps=3,
                       x_positions=np.array([xpos, xpos + 1e-7]), substrate_thickness=0.0)
    for i in range(3):
        zf = i / 3
        dc0 = (1 - zf) * og.dc0[0] + zf * og.dc0[1]
        dc1 = (1 - zf) * og.dc1[0] + zf * og.dc1[1]
        g1 = ravesim.DBAGrating(pitch0=4.2e-6, pitch1=12.6e-6, dc0=(dc0, dc0), dc1=(dc1, dc1),
                                z_start=0.0, thickness=t, nr_steps=1,
                                x_positions=[xpos, xpos + 1e-7], db_a=complex(0, beta))
        f = ravesim.Field(ctx, N, dx, dtype=dtype, chunk_size=chunk, plain=plain)
        f.upload(np.ones(N))
        f.apply_dba_grating(wl, g1, phase_step=1)
        got = np.abs(f.download()) < 0.75
        f.close()
        exp = np.concatenate([bw.dba_grating_mask(dx, k - N // 2, min(chunk, N - k), og, zf, 1)
                              for k in range(0, N, chunk)])
        assert got.shape == exp.shape
        assert np.array_equal(got, exp), f"slice {i}"


@pytest.mark.parametrize("dtype,tol,chunk", [("c16", 1e-10, None), ("c16", 1e-10, 3000),
                                             ("c8", 1e-4, None)])
def test_dba_grating_simdir_matches_oracle(ctx, tmp_path, dtype, tol, chunk):
    """A phase-stepped DBA grating with substrate through the sim-dir entry
    point (fast-wave's `dba_grating` element) against the oracle."""
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    from simcases import dba_case
    cfg, subs, ang = dba_case(dtype=dtype, chunk=chunk)
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    assert got.shape == ref.shape == (2, 40)
    err = rel_l2(got, ref)
    print(f"DBA grating {dtype} chunk {chunk} rel-L2 = {err:.3e}")
    assert err < tol
    assert rel_l2(got[0], got[1]) > 1e-3  # the two grating positions differ


@pytest.mark.parametrize("case,dtype,tol", [("sample", "c16", 1e-10), ("sample", "c8", 1e-4),
                                            ("dba", "c16", 1e-10), ("vector", "c16", 1e-10)])
def test_history_and_keypoints_match_oracle(ctx, tmp_path, case, dtype, tol):
    """history_dz and save_keypoints_path of run_single_simulation: the
    detector rows are unchanged, history.npy is (n_px, entries) with the
    oracle's z values exactly, keypoint files hold the field before and after
    every element of the first pass."""
    import shutil
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    from simcases import dba_case, sample_case
    if case == "sample":
        cfg, subs, ang = sample_case(tmp_path, dtype=dtype)
    else:
        cfg, subs, ang = dba_case(dtype=dtype)
    if case == "vector":  # a plane wave from 4 mm before the first element
        N = cfg["sim_params"]["N"]
        np.save(tmp_path / "ones.npy", np.ones(N, dtype=np.complex128))
        subs[0]["source"] = {"type": "vector", "input_path": str(tmp_path / "ones.npy"), "z": 0.006}
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    if case == "sample":
        shutil.copy(tmp_path / "sample_grid.npy", sd / "sample_grid.npy")
    hdz = 0.0037
    kdir = tmp_path / "kp"
    kdir.mkdir()
    got = run_single_simulation(sd, 0, tmp_path, save_keypoints_path=kdir, history_dz=hdz, ctx=ctx)
    kp = {}
    ref, h = bw.run_simdir(sd, 0, dtype=np.complex128, history_dz=hdz, keypoints=kp)
    assert rel_l2(got, ref) < tol
    sub = sd / "00000000"
    hist, hz, hx = (np.load(sub / f) for f in ("history.npy", "history_z.npy", "history_x.npy"))
    href = h.get_history()
    print(f"history {case} {dtype}: {hist.shape} rel-L2 {rel_l2(hist, href):.3e}")
    assert hist.shape == href.shape and hist.shape[1] > 8
    assert np.array_equal(hz, np.asarray(h.zs, dtype=np.float64))
    assert np.array_equal(hx, bw.detector_x_vector(cfg["sim_params"]["detector_size"],
                                                   cfg["sim_params"]["detector_pixel_size_x"]))
    assert rel_l2(hist, href) < tol
    for k in range(hist.shape[1]):
        if np.linalg.norm(href[:, k]) > 0:
            assert rel_l2(hist[:, k], href[:, k]) < 10 * tol, f"entry {k} z {hz[k]}"
    names = sorted(p.name for p in kdir.iterdir())
    assert names == sorted(f"keypoint_{i:02}_{s}.npy" for i, s in kp)
    ftol = 1e-10 if dtype == "c16" else 2e-4
    for (i, s), u in kp.items():
        assert rel_l2(np.load(kdir / f"keypoint_{i:02}_{s}.npy"), u) < ftol, f"keypoint {i} {s}"


def _load_energy(sim_dir):
    import yaml
    return float(yaml.safe_load((sim_dir / "00000000" / "subconfig.yaml").read_text())["energy"])


def test_save_and_exit_then_continue_matches_oracle(ctx, tmp_path):
    """big-wave/test.py:710-805: a base simulation cut at a save_and_exit
    element writes u_0000.npy there; a second simulation continues from it
    (vectors multisource).  Both against the oracle."""
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import setup_simulation
    from test_setup import DB, _config, _grating
    els = [_grating(0.1), {"type": "save_and_exit", "z_start": 0.25}, _grating(0.25), _grating(0.5)]
    cfg = _config(els)
    base = setup_simulation(cfg, tmp_path / "base", DB)
    got_b = run_single_simulation(base, 0, tmp_path, ctx=ctx)
    fu = []
    ref_b = bw.run_simdir(base, 0, final_u=fu)
    u = np.load(base / "00000000" / "u_0000.npy")
    assert u.dtype == np.complex128 and u.shape == (4096 * 4,)
    assert rel_l2(got_b, ref_b) < 1e-10
    # The global phase 2 pi z / wl is ~6e10 rad at z = 0.25 m; FP64 places it
    # to a few ulp of that (the oracle's own exp(-2 pi i dz / wl) too).  The
    # field agrees to 1e-10 once that constant is factored out, and the
    # constant lies within the rounding bound.
    c = np.vdot(fu[0], u)
    c /= abs(c)
    wl = 6.62607004e-34 * 299792458 / (_load_energy(base) * 1.602176634e-19)
    assert abs(np.angle(c)) < 8 * 2 * np.pi * 0.25 / wl * 2.0 ** -53
    assert rel_l2(u * np.conj(c), fu[0]) < 1e-10
    cont = dict(cfg, multisource={"type": "vectors", "base_sim_dir": str(base), "input_u_index": 0},
                elements=els[2:])
    sd = setup_simulation(cont, tmp_path / "cont", DB)
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = bw.run_simdir(sd, 0)
    print(f"continuation rel-L2 {rel_l2(got, ref):.3e}")
    assert got.shape == ref.shape == (1, 10)
    assert rel_l2(got, ref) < 1e-10


def test_fastwave_cli_matches_oracle(tmp_path):
    """The fastwave executable on a sim dir (phase-stepped DBA + grating
    chain), with --history_dz: detected.npy and history.npy against the
    oracle, files in big-wave's layout."""
    import subprocess
    from pathlib import Path
    from ravesim.simdir import write_simdir
    from simcases import dba_case
    exe = Path(__file__).resolve().parent.parent / "rave-sim_amd" / "ravesim" / "fastwave"
    cfg, subs, ang = dba_case(dtype="c16")
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    r = subprocess.run([str(exe), str(sd), "-s", "0", "--history_dz", "0.0037"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.load(sd / "00000000" / "detected.npy")
    ref, h = bw.run_simdir(sd, 0, dtype=np.complex128, history_dz=0.0037)
    assert got.dtype == np.float64 and got.shape == ref.shape
    assert rel_l2(got, ref) < 1e-10
    hist = np.load(sd / "00000000" / "history.npy")
    assert hist.shape == h.get_history().shape and rel_l2(hist, h.get_history()) < 1e-10


def test_fastwave_cli_tuning_table(tmp_path):
    """fastwave --tune: a table fixing the case's split (9|7), the shipped
    default and `none` all give rows that match the oracle; a table that
    cannot be read is an error (exit 1) with a message."""
    import subprocess
    from pathlib import Path
    import ravesim
    from ravesim.simdir import write_simdir
    from simcases import dba_case
    exe = Path(__file__).resolve().parent.parent / "rave-sim_amd" / "ravesim" / "fastwave"
    cfg, subs, ang = dba_case(dtype="c16")
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    ref = bw.run_simdir(sd, 0, dtype=np.complex128)
    table = tmp_path / "split.txt"
    c = ravesim.Context(0)
    set_plan_tiles(c, table, 2 ** 16, np.complex128, 9, 7, 0, 0)
    c.close()
    for extra in (["--tune", str(table)], [], ["--tune", "none"]):
        r = subprocess.run([str(exe), str(sd), "-s", "0", *extra], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (extra, r.stderr)
        assert rel_l2(np.load(sd / "00000000" / "detected.npy"), ref) < 1e-10, extra
    r = subprocess.run([str(exe), str(sd), "-s", "0", "--tune", str(tmp_path / "missing.txt")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "cannot load tuning table" in r.stderr


def test_default_context_loads_the_shipped_table():
    """Calls that pass no context use one with the shipped MI355X table: it
    holds the same entry for a 2^26 complex64 simulation as a fresh context
    that loads the table (none, on a device the table was not measured on)."""
    import ravesim
    from ravesim import multisim
    fresh = ravesim.Context(0)
    n = fresh.tune_load(multisim.DEFAULT_TUNE_TABLE)
    expect = fresh.tuned(2 ** 26, np.complex64, 1)
    fresh.close()
    assert multisim._ctx().tuned(2 ** 26, np.complex64, 1) == expect
    assert (expect is not None) == (n > 0)


def test_talbot_lau_full_size_precision_form(ctx, tmp_path):
    """The precision form at the benchmark's full size: Talbot-Lau chain at
    N = 2^26 with G2 stepped over 8 positions (16 slices per grating).  Beyond
    the oracle's reach the reference is the plain complex128 GPU form (itself
    checked against the oracle at smaller N); the complex64 + FP64-phase form
    stays within 1e-4 relative L2 of intensity and 1e-3 rad of fringe phase."""
    from ravesim.multisim import run_single_simulation
    sd = _simdir(tmp_path, talbot_lau(tmp_path, N=2 ** 26, slices=16, det_size=1e-3, dtype="c8", g2_steps=8))
    got = run_single_simulation(sd, 0, tmp_path, ctx=ctx)
    ref = run_single_simulation(sd, 0, tmp_path, ctx=ctx, dtype="c16", plain=True)
    assert got.shape == ref.shape == (8, 1000)
    err = rel_l2(got, ref)
    k = np.arange(8)
    w = np.exp(2j * np.pi * k / 8)[:, None]
    a, b = (w * got).sum(0), (w * ref).sum(0)
    ok = np.abs(b) > 1e-3 * np.abs(ref).mean() * 8
    dphi = np.abs(np.angle(a[ok] * np.conj(b[ok]))).max()
    print(f"Talbot-Lau N=2^26 c8 vs plain c16: rel-L2 {err:.3e}, fringe phase {dphi:.2e} rad over {ok.sum()} px")
    assert err < 1e-4 and dphi < 1e-3


def test_n30_field_round_trip_and_composition(ctx):
    """A 2^30-sample complex64 field (8 GiB, the largest served size): a point
    source generated on the device, propagated forward and back, returns to
    its detector image; two hops equal one."""
    import ravesim
    N, dx, wl = 2 ** 30, 2e-9, 2.695e-11
    size, px = 1e-3, 1e-6
    a = ravesim.Field(ctx, N, dx, dtype=np.complex64)
    a.point_source(wl, 1e-7, 200.0)  # Nyquist at 2^30 samples needs z > 160 m
    d0 = a.square_and_downsample(size, px, 1.0, 200.0)
    a.propagate(wl, 0.05)
    a.propagate(wl, 0.07)
    d1 = a.square_and_downsample(size, px, 1.0, 200.12)
    a.propagate(wl, -0.12)
    d2 = a.square_and_downsample(size, px, 1.0, 200.0)
    a.close()
    b = ravesim.Field(ctx, N, dx, dtype=np.complex64)
    b.point_source(wl, 1e-7, 200.0)
    b.propagate(wl, 0.12)
    d3 = b.square_and_downsample(size, px, 1.0, 200.12)
    b.close()
    print(f"2^30: round trip {rel_l2(d2, d0):.2e}, two hops vs one {rel_l2(d1, d3):.2e}")
    assert np.all(d0 > 0) and rel_l2(d2, d0) < 1e-5 and rel_l2(d1, d3) < 1e-5


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-10), (np.complex64, 2e-5)])
def test_sample_operator_matches_oracle(ctx, dtype, tol):
    """rs_apply_sample (fast-wave's apply_sample_factors + propagation per
    grid row) against the oracle's Sample.apply, at the second x position and
    with a cutoff."""
    import ravesim
    N, dx, wl = 2 ** 16, 1e-9, 4.1e-11
    rng = np.random.default_rng(4)
    grid = rng.integers(0, 3, size=(5, 24)).astype(np.int8)
    dbl = [complex(5.1e-7, 1.6e-10), complex(3.6e-6, 2.1e-7)]
    xs = [0.0, 1.7e-7]
    fc = 0.35 / dx
    u = rand_field(N, 9)
    f = ravesim.Field(ctx, N, dx, dtype=dtype)
    f.upload(u)
    f.apply_sample(wl, ravesim.Sample(z_start=1e-3, pixel_size_x=4e-7, pixel_size_z=1.5e-6, grid=grid,
                                      deltabeta=dbl, x_positions=xs), phase_step=1, cutoff_freq=fc)
    got = f.download()
    f.close()
    p = bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                     detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
    ref = bw.Fields(u.astype(np.complex128))
    s = bw.Sample(z_start=1e-3, pixel_size_x=4e-7, pixel_size_z=1.5e-6, grid=grid,
                  db_list=np.array([0] + dbl, dtype=np.complex128), x_positions=np.array(xs))
    bw.sample_apply(ref, s, p, fc, 1)
    assert rel_l2(got, ref.u) < tol


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-10), (np.complex64, 1e-4)])
def test_batch_sources_summed_per_output_row(ctx, dtype, tol):
    """Point sources of several simulations in one lockstep batch, each
    simulation's intensities summed into its own detector row (out_index),
    against the oracle's per-source runs summed incoherently."""
    import ravesim
    from ravesim.core import BatchMember, run_batch
    N, dx, zs = 2 ** 16, 2e-9, -5.0
    sims = [(30.0, 3.1e-6, 0.45, [0.2e-6]), (45.0, 2.2e-6, 0.55, [-0.4e-6, 0.1e-6, 0.35e-6]),
            (62.0, 5.3e-6, 0.5, [-0.1e-6, 0.45e-6])]
    h, slices, z_det = 30e-6, 8, 30e-6 + 0.1
    ms, ref = [], []
    for r, (e, pitch, dc, xs) in enumerate(sims):
        sc = 25.0 / e
        db = complex(7.7e-7 * sc ** 2, 3.5e-10 * sc ** 4)
        wl = ravesim.energy_to_wavelength(e * 1000.0)
        g = ravesim.Grating(pitch=pitch, dc=(dc, dc), z_start=0.0, thickness=h, nr_steps=slices, db_a=db)
        og = bw.Grating(pitch=pitch, dc=(dc, dc), z_start=0.0, thickness=h, nr_steps=slices,
                        x_positions=np.array([0.0]), substrate_thickness=0.0, db_a=db)
        p = bw.SimParams(N=N, dx=dx, z_detector=z_det, detector_size=N * dx / 4, detector_pixel_size_x=1e-6,
                         detector_pixel_size_y=1.0, wl=wl, chunk_size=N)
        acc = 0
        for x in xs:
            ms.append(BatchMember(wl=wl, gratings=[g], cutoff_angles=[math.pi / 2] * 2, out_index=r,
                                  source_x=x, source_z=zs))
            acc = acc + bw.run_simulation(p, bw.PointSource(x=x, z=zs), [og], [math.pi / 2] * 2)[0]
        ref.append(acc)
    got = run_batch(ctx, N, dx, ms, n_outputs=3, dtype=dtype, det_size=N * dx / 4, det_px=1e-6, det_py=1.0,
                    z_detector=z_det)
    ref = np.array(ref)
    assert got.shape == ref.shape == (3, 32)
    for r in range(3):
        assert rel_l2(got[r], ref[r]) < tol, f"simulation {r}"


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-10), ("c8", 1e-4)])
@pytest.mark.parametrize("seed", range(16))
def test_random_configs_match_oracle(ctx, tmp_path, seed, dtype, tol):
    """Seeded random sim dirs (gratings, DBA gratings and samples in random
    order with random substrates, chunk sizes, cutoffs, point or plane-wave
    sources, one element phase stepped) with history: detector rows and
    history against the complex128 oracle, z values bit-equal."""
    from ravesim.multisim import run_single_simulation
    from ravesim.simdir import write_simdir
    from simcases import random_case
    cfg, subs, ang, grid = random_case(tmp_path, seed)
    cfg["dtype"] = dtype
    sd = write_simdir(tmp_path / "sim", cfg, subs, ang)
    if grid is not None:
        np.save(sd / "grid.npy", grid)
    got = run_single_simulation(sd, 0, tmp_path, history_dz=0.003, ctx=ctx)
    ref, h = bw.run_simdir(sd, 0, dtype=np.complex128, history_dz=0.003)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < tol
    hist = np.load(sd / "00000000" / "history.npy")
    assert np.array_equal(np.load(sd / "00000000" / "history_z.npy"), np.asarray(h.zs, dtype=np.float64))
    assert rel_l2(hist, h.get_history()) < tol


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-10), (np.complex64, 1e-4)])
@pytest.mark.parametrize("seed", range(8))
def test_random_batches_match_oracle(ctx, seed, dtype, tol):
    """Seeded random lockstep batches: one structure per batch (source kind,
    gratings, slice counts, substrates), per-member energies, pitches, duty
    cycles, positions, thicknesses, materials and cutoffs, random weights and
    output rows, random lanes and group sizes; against the oracle's weighted
    per-member sums."""
    from ravesim.core import BatchMember, run_batch
    import ravesim
    rng = np.random.default_rng(100 + seed)
    N = int(2 ** rng.integers(12, 16))
    dx = 1e-9
    chunk = int(rng.choice([N, 1000]))
    ng = int(rng.integers(1, 3))
    steps = [int(rng.integers(1, 4)) for _ in range(ng)]
    has_sub = [bool(rng.random() < 0.5) for _ in range(ng)]
    z0 = [0.01 + 0.01 * k for k in range(ng)]
    z_det = z0[-1] + 0.01 + float(rng.uniform(0, 0.02))
    point = bool(rng.random() < 0.5)
    n_out = int(rng.integers(1, 4))
    det, px = min(N * dx / 2, 1e-5), 2e-7
    ms, ref = [], np.zeros((n_out, int(det // px)))
    for _ in range(int(rng.integers(2, 9))):
        e = float(rng.uniform(20e3, 60e3))
        wl = ravesim.energy_to_wavelength(e)
        gs, ogs = [], []
        for k in range(ng):
            pitch = float(rng.uniform(0.3e-6, 1.5e-6))
            dc = (float(rng.uniform(0.3, 0.7)), float(rng.uniform(0.3, 0.7)))
            xs = [float(rng.uniform(-pitch, pitch))]
            th = float(rng.uniform(1e-6, 5e-6))
            sub = float(rng.uniform(1e-6, 3e-6)) if has_sub[k] else 0.0
            dba, dbb, dbs = (complex(rng.uniform(1e-7, 4e-6), rng.uniform(0, 2e-7)) for _ in range(3))
            gs.append(ravesim.Grating(pitch=pitch, dc=dc, z_start=z0[k], thickness=th, nr_steps=steps[k],
                                      x_positions=xs, substrate_thickness=sub, db_a=dba, db_b=dbb,
                                      db_substrate=dbs))
            ogs.append(bw.Grating(pitch=pitch, dc=dc, z_start=z0[k], thickness=th, nr_steps=steps[k],
                                  x_positions=np.array(xs), substrate_thickness=sub, db_a=dba, db_b=dbb,
                                  db_substrate=dbs))
        ang = sorted(float(rng.choice([float(rng.uniform(0.2e-3, 2e-3)), math.pi / 2])) for _ in range(ng + 1))
        w, r = float(rng.uniform(0.1, 1.0)), int(rng.integers(0, n_out))
        xsrc = float(rng.uniform(-1e-7, 1e-7))
        ms.append(BatchMember(wl=wl, gratings=gs, cutoff_angles=ang, weight=w, out_index=r,
                              plane_wave=not point, source_x=xsrc if point else 0.0,
                              source_z=-0.5 if point else z0[0]))
        p = bw.SimParams(N=N, dx=dx, z_detector=z_det, detector_size=det, detector_pixel_size_x=px,
                         detector_pixel_size_y=1.0, wl=wl, chunk_size=chunk)
        src = (bw.PointSource(x=xsrc, z=-0.5) if point
               else bw.VectorSource(vec=np.ones(N, dtype=np.complex128), z=z0[0]))
        ref[r] += w * bw.run_simulation(p, src, ogs, ang)[0]
    got = run_batch(ctx, N, dx, ms, n_outputs=n_out, dtype=dtype, chunk_size=chunk, det_size=det, det_px=px,
                    det_py=1.0, z_detector=z_det, lanes=int(rng.integers(0, 4)),
                    max_resident=int(rng.integers(0, 4)))
    assert got.shape == ref.shape
    for k in range(n_out):
        if np.any(ref[k]):
            assert rel_l2(got[k], ref[k]) < tol, f"row {k}"
        else:
            assert not np.any(got[k])


@pytest.mark.parametrize("seed", range(12))
def test_random_splits_match_oracle(seed):
    """Random N and random valid digit-group splits (two or three groups,
    wide columns included): a grating (slices, substrate, multi-chunk x) and a
    propagation with cutoff through the operators, against the oracle."""
    import ravesim
    from ravesim import _lib
    rng = np.random.default_rng(200 + seed)
    dtype = np.complex128 if seed % 2 else np.complex64
    f64 = dtype == np.complex128
    lgN = int(rng.integers(10, 21))
    N, dx, wl = 2 ** lgN, 2e-9, 2.7e-11
    c = ravesim.Context(0)
    split = None
    for _ in range(50):
        r = int(rng.integers(6, (12 if f64 else 13) + 1))
        a = int(rng.integers(4, (12 if f64 else 13) + 1))
        try:
            c.set_split(N, dtype, r, a)
        except _lib.RsError:
            continue
        split = (r, a)
        break
    assert split is not None
    chunk = int(rng.choice([N, max(N // 3, 64)]))
    u = rand_field(N, seed)
    pitch = float(rng.uniform(0.5e-6, 3e-6))
    g = ravesim.Grating(pitch=pitch, dc=(0.35, 0.6), z_start=0.0, thickness=6e-6, nr_steps=3,
                        x_positions=[1.1e-7], substrate_thickness=2e-6, db_a=complex(3e-6, 1e-7),
                        db_substrate=complex(5e-7, 1e-9))
    fc = float(rng.uniform(0.2, 0.5)) / dx
    f = ravesim.Field(c, N, dx, dtype=dtype, chunk_size=chunk)
    f.upload(u)
    f.apply_grating(wl, g, cutoff_freq=fc)
    f.propagate(wl, 1.3e-3, fc)
    got = f.download()
    f.close()
    c.close()
    p = bw.SimParams(N=N, dx=dx, z_detector=1.0, detector_size=1e-6, detector_pixel_size_x=1e-7,
                     detector_pixel_size_y=1.0, wl=wl, chunk_size=chunk)
    ref = bw.Fields(u.astype(np.complex128))
    og = bw.Grating(pitch=pitch, dc=(0.35, 0.6), z_start=0.0, thickness=6e-6, nr_steps=3,
                    x_positions=np.array([1.1e-7]), substrate_thickness=2e-6, db_a=complex(3e-6, 1e-7),
                    db_substrate=complex(5e-7, 1e-9))
    bw.grating_apply(ref, og, p, fc, 0)
    bw.propagate(ref, dx, wl, 1.3e-3, chunk, fc)
    err = rel_l2(got, ref.u)
    print(f"2^{lgN} {'c16' if f64 else 'c8'} split {split} chunk {chunk}: {err:.2e}")
    assert err < (1e-10 if f64 else 2e-5)
