"""Pins the CPU oracle (oracle/tdc_oracle.c) without trusting its own formulas.

Each check is chosen so that a plausible mistake fails it:
  * the reference's known answers for the math helpers (tdctest/mathtest.cpp),
  * scipy's window implementations,
  * closed forms: a point target focuses to a * sum(w) only with the right
    phase sign and scale; Doppler frequencies for velocities perpendicular /
    parallel to the line of sight; window values at the centre, the band edge
    and outside; the skip test on geometries that decide it,
  * invariants: linearity, superposition, permuting chirps with their
    trajectory,
  * an arbitrary-precision (mpmath) brute force on a tiny scene,
  * the FP32 reference-method restatement agreeing to FP32 accuracy.
"""
import json
import os

import numpy as np
import pytest
import scipy.signal.windows as sw

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def vectors():
    with open(os.path.join(HERE, "golden", "mathtest_vectors.json")) as fh:
        return json.load(fh)


# ---------------------------------------------------------------------------- helpers


def test_norm3_known_answer(orc, vectors):
    for v in vectors["norm3"]:
        assert orc.norm3(*v["in"]) == pytest.approx(v["out"], rel=1e-15)


def test_dot3_stated_values(orc, vectors):
    for v in vectors["dot3"]:
        assert orc.dot3(v["a"], v["b"]) == v["out"]


def test_q_conj_known_answer(orc, vectors):
    for v in vectors["q_conj"]:
        assert np.array_equal(orc.q_conj(v["q"]), v["out"])


def test_q_prod_known_answer(orc, vectors):
    for v in vectors["q_prod"]:
        assert np.array_equal(orc.q_prod(v["a"], v["b"]), v["out"])


def test_q_rot_known_answer(orc, vectors):
    for v in vectors["q_rot"]:
        assert np.allclose(orc.q_rot(v["q"], v["v"]), v["out"], atol=1e-7)


def test_q_rot_invariants(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        v = rng.normal(size=3)
        r = orc.q_rot(q, v)
        assert np.linalg.norm(r) == pytest.approx(np.linalg.norm(v), rel=1e-12)
        # q (x) q* = |q|^2 (scalar), and rotating by q then by q* is the identity
        assert np.allclose(orc.q_prod(q, orc.q_conj(q)), [1, 0, 0, 0], atol=1e-12)
        assert np.allclose(orc.q_rot(orc.q_conj(q), r), v, atol=1e-12)
    # a rotation about z by angle a: scalar cos(a/2), vector (0, 0, sin(a/2))
    a = 0.3
    r = orc.q_rot([np.cos(a / 2), 0, 0, np.sin(a / 2)], [1.0, 0.0, 0.0])
    assert np.allclose(r, [np.cos(a), np.sin(a), 0.0], atol=1e-15)


def test_float_pair_arithmetic_vectors(vectors):
    # float2 + and float3 - of the reference restated on float32 values
    for v in vectors["float2_add"]:
        assert np.array_equal(np.float32(v["a"]) + np.float32(v["b"]), np.float32(v["out"]))
    for v in vectors["float3_sub"]:
        assert np.array_equal(np.float32(v["a"]) - np.float32(v["b"]), np.float32(v["out"]))


# ---------------------------------------------------------------------------- windows


def test_range_window_single_sample(orc):
    # the reference formula at Ns = 1 is A - (1 - A) cos 0 = 2A - 1 (scipy special-cases it)
    assert orc.range_window(1) == pytest.approx([2 * 25 / 46 - 1], abs=1e-15)


@pytest.mark.parametrize("n", [2, 7, 8, 100, 128, 512])
def test_range_window_matches_scipy(orc, n):
    assert np.allclose(orc.range_window(n), sw.general_hamming(n, 25 / 46, sym=False), atol=1e-14)
    assert np.allclose(orc.range_window(n, 0.5), sw.hann(n, sym=False), atol=1e-14)
    assert np.array_equal(orc.range_window(n, apply=False), np.ones(n))


@pytest.mark.parametrize("n,nbar,sll", [(8, 4, 30.0), (100, 4, 30.0), (512, 4, 30.0),
                                        (64, 6, 40.0), (33, 3, 25.0)])
def test_taylor_window_matches_scipy(orc, n, nbar, sll):
    assert np.allclose(orc.taylor_window(n, nbar, sll), sw.taylor(n, nbar, sll, sym=False),
                       atol=1e-12)
    assert np.allclose(orc.taylor_window(n, nbar, sll, periodic=False),
                       sw.taylor(n, nbar, sll, sym=True), atol=1e-12)


def test_scene_windows_match_scipy():
    from sarlabtdc import scene

    assert np.allclose(scene.hann(64), sw.hann(64, sym=False), atol=1e-15)
    assert np.allclose(scene.taylor(512, 4, 30.0), sw.taylor(512, 4, 30.0, sym=False), atol=1e-12)


# ---------------------------------------------------------------------------- Doppler closed forms


def test_doppler_frequency_closed_forms(orc):
    lam = 3.9e-3
    pos = np.array([0.0, 0.0, 0.0])
    tgt = np.array([0.0, 2.0, 0.0])
    # velocity perpendicular to the line of sight: zero Doppler
    assert orc.doppler_freq(pos, [5.0, 0.0, 0.0], tgt, lam) == 0.0
    # parallel: +2|v|/lambda towards the target, -2|v|/lambda away from it
    assert orc.doppler_freq(pos, [0.0, 5.0, 0.0], tgt, lam) == pytest.approx(2 * 5 / lam, rel=1e-15)
    assert orc.doppler_freq(pos, [0.0, -5.0, 0.0], tgt, lam) == pytest.approx(-2 * 5 / lam, rel=1e-15)
    # 45 degrees: 2 |v| cos(45)/lambda
    f = orc.doppler_freq(pos, [5.0, 0.0, 0.0], [2.0, 2.0, 0.0], lam)
    assert f == pytest.approx(2 * 5 * np.sqrt(0.5) / lam, rel=1e-14)
    # target at the antenna: 0/0
    assert np.isnan(orc.doppler_freq(pos, [1.0, 0.0, 0.0], pos, lam))
    # centroid with an identity attitude: 2/lambda v . boresight
    ant = orc.q_rot([1, 0, 0, 0], [0, 1, 0])
    assert orc.doppler_centroid([3.0, 4.0, 0.0], ant, lam) == pytest.approx(8.0 / lam, rel=1e-15)


def test_doppler_window_values(orc):
    a = 25 / 46
    bw = 1000.0
    assert orc.doppler_window(10.0, 10.0, bw, True) == pytest.approx(1.0, abs=1e-15)
    assert orc.doppler_window(500.0, 0.0, bw, True) == pytest.approx(2 * a - 1, abs=1e-12)
    assert orc.doppler_window(-500.0, 0.0, bw, True) == pytest.approx(2 * a - 1, abs=1e-12)
    assert orc.doppler_window(500.001, 0.0, bw, True) == 0.0
    assert orc.doppler_window(250.0, 0.0, bw, True) == pytest.approx(a, abs=1e-12)
    assert orc.doppler_window(499.0, 0.0, bw, False) == 1.0
    assert orc.doppler_window(501.0, 0.0, bw, False) == 0.0
    assert orc.doppler_window(float("nan"), 0.0, bw, False) == 0.0


def test_center_wavelength(orc):
    # c / ((f0 + f0 + K T) / 2) for f0 = 77 GHz, K T = 4 GHz -> 79 GHz
    lam = orc.center_wavelength(0.0, 40e-6, 1e14, 77e9)
    assert lam == pytest.approx(299792458.0 / 79e9, rel=1e-15)


# ---------------------------------------------------------------------------- image closed forms


def _small_scene(nc=24, ns=32, seed=1):
    from sarlabtdc import scene

    rng = np.random.default_rng(seed)
    track = np.zeros((nc, 3))
    track[:, 0] = np.linspace(-0.1, 0.1, nc)
    track += rng.normal(scale=1e-4, size=track.shape)
    t = np.arange(ns) * (scene.CHIRP_TIME / ns)
    grid = scene.plane_grid(6, 7, 4e-3, (0.0, 0.4, 0.0))
    return track, t, grid


def test_point_target_focuses_to_amplitude_times_window_sum(orc):
    from sarlabtdc import scene

    track, t, grid = _small_scene()
    p0 = grid[2, 3]
    a = 0.6 + 0.8j
    raw = scene.simulate_beat(track, t, p0[None], [a])
    rw = orc.range_window(t.size)
    img = orc.image_f64(raw, t, track, grid, scene.MOD_RATE, scene.F0, range_win=rw)
    want = a * track.shape[0] * rw.sum()
    assert abs(img[2, 3] - want) / abs(want) < 1e-12
    # the focus is the brightest pixel, and the phase sign matters: the
    # conjugate scene (wrong sign) does not focus
    assert np.unravel_index(np.argmax(np.abs(img)), img.shape) == (2, 3)
    img_conj = orc.image_f64(np.conj(raw), t, track, grid, scene.MOD_RATE, scene.F0)
    assert np.abs(img_conj).max() < 0.2 * abs(want)


def test_point_target_per_sample_positions(orc):
    """Per-sample positions (the reference layout) equal per-chirp ones when broadcast."""
    from sarlabtdc import scene

    track, t, grid = _small_scene()
    raw = scene.simulate_beat(track, t, grid[1, 1][None], [1.0])
    a = orc.image_f64(raw, t, track, grid, scene.MOD_RATE, scene.F0)
    ps = np.broadcast_to(track[:, None, :], (track.shape[0], t.size, 3))
    b = orc.image_f64(raw, t, ps, grid, scene.MOD_RATE, scene.F0)
    assert np.allclose(a, b, rtol=0, atol=1e-9)


def test_linearity_superposition_and_permutation(orc):
    from sarlabtdc import scene

    track, t, grid = _small_scene()
    rng = np.random.default_rng(2)
    r1 = rng.normal(size=(track.shape[0], t.size)) + 1j * rng.normal(size=(track.shape[0], t.size))
    r2 = scene.simulate_beat(track, t, grid[4, 5][None], [2.0 - 1j])
    f = lambda r, tr=track: orc.image_f64(r, t, tr, grid, scene.MOD_RATE, scene.F0)
    i1, i2 = f(r1), f(r2)
    assert np.allclose(f(3.0 * r1 - 2.0 * r2), 3.0 * i1 - 2.0 * i2, rtol=0,
                       atol=1e-10 * np.abs(i1).max())
    perm = rng.permutation(track.shape[0])
    assert np.allclose(f(r1[perm], track[perm]), i1, rtol=0, atol=1e-9 * np.abs(i1).max())


def test_mpmath_brute_force_with_doppler(orc):
    """Independent arbitrary-precision evaluation of the reference formula."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    rng = np.random.default_rng(4)
    nc, ns = 6, 5
    t = np.arange(ns) * 8e-6
    pos = rng.normal(scale=0.05, size=(nc, ns, 3))
    vel = rng.normal(scale=3.0, size=(nc, ns, 3))
    att = rng.normal(size=(nc, ns, 4))
    att /= np.linalg.norm(att, axis=-1, keepdims=True)
    raw = rng.normal(size=(nc, ns)) + 1j * rng.normal(size=(nc, ns))
    grid = rng.normal(scale=0.05, size=(3, 3)) + np.array([0.0, 0.5, 0.0])
    bore = np.array([0.0, 1.0, 0.0])
    k_rate, f0 = 1e14, 77e9
    lam = orc.center_wavelength(t[0], t[-1], k_rate, f0)
    rw = orc.range_window(ns)
    bw = 2.2e3
    dop = dict(velocity=vel, attitude=att, boresight=bore, center=0.0, bandwidth=bw,
               centroid=True, taper=True, lambda_c=lam)
    got = orc.image_f64(raw, t, pos, grid, k_rate, f0, range_win=rw, doppler=dop)

    mp = mpmath.mpf
    A = mp(25) / 46

    def qmul(a, b):
        return [a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]]

    want = []
    for p in grid:
        acc = mpmath.mpc(0)
        for n in range(nc):
            for m in range(ns):
                q = [mp(x) for x in att[n, m]]
                r = qmul(qmul(q, [mp(0)] + [mp(x) for x in bore]), [q[0], -q[1], -q[2], -q[3]])
                ant = r[1:]
                v = [mp(x) for x in vel[n, m]]
                d = [mp(p[i]) - mp(pos[n, m, i]) for i in range(3)]
                rng_ = mpmath.sqrt(sum(x * x for x in d))
                fc = 2 / mp(lam) * sum(v[i] * ant[i] for i in range(3))
                fd = 2 / mp(lam) * sum(v[i] * d[i] for i in range(3)) / rng_
                delta = fd - fc
                if abs(delta) <= mp(bw) / 2:
                    wa = A - (1 - A) * mpmath.cos(2 * mpmath.pi * delta / bw - mpmath.pi)
                else:
                    wa = mp(0)
                f = mp(f0) + mp(k_rate) * mp(t[m])
                ph = 2 * mpmath.pi * 2 * f * rng_ / mp(299792458)
                acc += mpmath.mpc(raw[n, m].real, raw[n, m].imag) * mp(rw[m]) * wa * mpmath.expj(ph)
        want.append(complex(acc))
    want = np.array(want)
    # every chunk is processed here (one partial chunk, probes in band), so the
    # image is the windowed direct sum
    assert np.abs(want).max() > 0
    assert np.allclose(got, want, rtol=0, atol=1e-9 * np.abs(want).max())


# ---------------------------------------------------------------------------- skip heuristic


def _skip_scene(npri=140, ns=4):
    t = np.arange(ns) * 1e-5
    pos = np.zeros((npri, ns, 3))
    pos[:, :, 0] = np.linspace(-0.1, 0.1, npri)[:, None]
    vel = np.zeros((npri, ns, 3))
    att = np.zeros((npri, ns, 4))
    att[..., 0] = 1.0
    raw = np.ones((npri, ns), complex)
    return t, pos, vel, att, raw


def _dop(vel, att, bw=1.0):
    return dict(velocity=vel, attitude=att, boresight=[0, 1, 0], center=0.0, bandwidth=bw,
                centroid=False, taper=False, lambda_c=3.8e-3)


def test_skip_partial_chunk_reads_previous_rows(orc):
    """Probes beyond the valid rows of a partial chunk read the previous chunk's rows
    (tdcprocessor.cpp:325-345 and :355-365)."""
    t, pos, vel, att, raw = _skip_scene()
    target = np.array([0.0, 1.0, 0.0])
    # chunk 1 has rows 128..139; its valid rows move fast along the line of sight
    vel[128:, :, 1] = 50.0
    assert orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 0, target)
    # rows 63 and 127 of chunk 0 are at rest: probes (63, .) and (127, .) of
    # chunk 1 see them and keep the chunk
    assert orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 1, target)
    # move chunk-0 rows 63 and 127 as well: now every probe is out of band
    vel[63, :, 1] = 50.0
    vel[127, :, 1] = 50.0
    assert not orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 1, target)


def test_skip_first_partial_chunk_reads_zeros(orc):
    """A partial first chunk probes zero-filled rows: position 0, velocity 0."""
    t, pos, vel, att, raw = _skip_scene(npri=40)
    vel[:, :, 1] = 50.0
    target = np.array([0.0, 1.0, 0.0])
    # rows 63 and 127 do not exist: zero velocity there gives f_D = 0, in band
    assert orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 0, target)
    # a target at the origin makes those zero probes 0/0 (ignored: not first)
    # while the first probe is valid and, moving along x, out of band
    vel[:, :, 1] = 0.0
    vel[:, :, 0] = 50.0
    assert not orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 0,
                                   np.array([0.0, 0.0, 0.0]))


def test_skip_nan_first_probe_skips(orc):
    t, pos, vel, att, raw = _skip_scene()
    tgt = pos[0, 0].copy()   # the first probe of chunk 0: 0/0
    assert not orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att, bw=1e9), 0, tgt)
    tgt = pos[63, 0].copy()  # a later probe: ignored by the minimum
    assert orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att, bw=1e9), 0, tgt)


def test_skip_probe_samples(orc):
    """Sample probes are max(j*(ns/2)-1, 0): {0, ns/2-1, 2*(ns/2)-1}."""
    t, pos, vel, att, raw = _skip_scene(ns=5)
    target = np.array([0.0, 1.0, 0.0])
    vel[:, :, 1] = 50.0
    vel[:, 3, 1] = 0.0   # sample 3 = 2*(5//2)-1 is probed
    assert orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 0, target)
    vel[:, 3, 1] = 50.0
    vel[:, 4, 1] = 0.0   # sample 4 is never probed
    assert not orc.chunk_processed(raw, t, pos, 1e14, 77e9, _dop(vel, att), 0, target)


def test_skipped_chunk_contributes_nothing(orc):
    t, pos, vel, att, raw = _skip_scene()
    target = np.array([[0.0, 1.0, 0.0]])
    vel[128:, :, 1] = 50.0
    vel[63, :, 1] = 50.0
    vel[127, :, 1] = 50.0
    dop = _dop(vel, att, bw=1.0)
    img = orc.image_f64(raw, t, pos, target, 1e14, 77e9, doppler=dop)
    # chunk 1 skipped; in chunk 0 rows 63 and 127 are windowed out per term
    no_dop = orc.image_f64(raw[:128], t, pos[:128], target, 1e14, 77e9)
    moving = orc.image_f64(raw[[63, 127]], t, pos[[63, 127]], target, 1e14, 77e9)
    assert np.allclose(img, no_dop - moving, rtol=1e-12, atol=1e-9)


# ---------------------------------------------------------------------------- reference method


def test_reference_method_agrees_with_oracle_line_small(orc):
    from golden.make_golden import load_dropin

    d, z = load_dropin("line_small_plain.npz")
    rm = orc.refmethod_f32(d, 0.0, 1e30, False, False, False, doppler=False)
    assert orc.rel_l2(rm, z["image_f64"]) < 5e-4
    rm2 = orc.refmethod_f32(d, 0.0, 1e30, False, False, False, doppler=True)
    assert np.array_equal(rm, rm2)


def test_reference_method_agrees_with_oracle_doppler(orc):
    from golden.make_golden import load_dropin

    d, z = load_dropin("doppler_small.npz")
    rm = orc.refmethod_f32(d, 0.0, 8e3, True, True, True)
    assert orc.rel_l2(rm, z["image_f64"]) < 5e-3


# ---------------------------------------------------------------------------- golden fixtures


def test_golden_line_small_reproduces(orc):
    from golden.make_golden import load_dropin

    d, z = load_dropin("line_small_plain.npz")
    img = orc.image_f64(d["rawData"], d["sampleTimes"], d["position"], d["focusGrid"],
                        float(d["modRate"]), float(d["startFreq"]))
    assert np.allclose(img, z["image_f64"], rtol=0, atol=1e-9 * np.abs(img).max())


def test_golden_doppler_reproduces(orc):
    from golden.make_golden import load_dropin

    d, z = load_dropin("doppler_small.npz")
    dop = dict(velocity=d["velocity"], attitude=d["attitude"], boresight=d["bodyBoresight"],
               center=0.0, bandwidth=8e3, centroid=True, taper=True)
    img = orc.image_f64(d["rawData"], d["sampleTimes"], d["position"], d["focusGrid"],
                        float(d["modRate"]), float(d["startFreq"]),
                        range_win=orc.range_window(d["rawData"].shape[1]), doppler=dop)
    assert np.allclose(img, z["image_f64"], rtol=0, atol=1e-9 * np.abs(img).max())
    assert 0 < int(z["processed_chunks"]) < d["focusGrid"].shape[0] * d["focusGrid"].shape[1] * 3


def test_golden_precision_reproduces(orc):
    z = np.load(os.path.join(HERE, "golden", "precision_ragged.npz"))
    img = orc.image_f64(z["rawData"], z["sampleTimes"], z["position"], z["focusGrid"],
                        float(z["modRate"]), float(z["startFreq"]))
    assert np.allclose(img, z["image_f64"], rtol=0, atol=1e-9 * np.abs(img).max())
