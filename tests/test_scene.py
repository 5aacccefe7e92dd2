"""The beat-signal simulator against a direct numpy evaluation."""
import numpy as np


def direct(track, t, scat, amp, f0, k_rate, c=299792458.0):
    r = np.linalg.norm(track[:, None, :] - scat[None, :, :], axis=-1)      # [nc, nk]
    f = f0 + k_rate * t                                                    # [ns]
    ph = -2j * np.pi * 2.0 * f[None, :, None] * r[:, None, :] / c           # [nc, ns, nk]
    return np.sum(np.exp(ph) * amp[None, None, :], axis=-1)


def test_uniform_times_recurrence_matches_direct():
    from sarlabtdc import scene

    rng = np.random.default_rng(0)
    track = rng.normal(scale=0.3, size=(17, 3))
    scat = rng.normal(scale=0.5, size=(9, 3)) + [0, 2, 0]
    amp = rng.normal(size=9) + 1j * rng.normal(size=9)
    t = np.arange(200) * 40e-6 / 200
    got = scene.simulate_beat(track, t, scat, amp)
    want = direct(track, t, scat, amp, scene.F0, scene.MOD_RATE)
    assert np.abs(got - want).max() < 1e-9 * np.abs(amp).sum()


def test_nonuniform_times_direct_path():
    from sarlabtdc import scene

    rng = np.random.default_rng(1)
    track = rng.normal(scale=0.3, size=(5, 3))
    scat = rng.normal(scale=0.5, size=(4, 3)) + [0, 1, 0]
    amp = rng.normal(size=4) + 1j * rng.normal(size=4)
    t = np.sort(rng.uniform(0, 40e-6, size=33))
    got = scene.simulate_beat(track, t, scat, amp)
    want = direct(track, t, scat, amp, scene.F0, scene.MOD_RATE)
    assert np.abs(got - want).max() < 1e-11 * np.abs(amp).sum()


def test_configs_shapes_and_seeds():
    from sarlabtdc import scene

    a, b = scene.line_small(), scene.line_small()
    assert a.raw.shape == (64, 128) and a.grid.shape == (32, 32, 3)
    assert np.array_equal(a.raw, b.raw)
    assert not np.array_equal(a.raw, scene.line_small(seed=1).raw)
    d = a.dropin()
    assert d["position"].shape == (64, 128, 3) and d["position"].dtype == np.float32
    assert d["attitude"].shape == (64, 128, 4)
    shapes = scene.stream_shapes(1000)
    nc = np.array([s[0] for s in shapes])
    side = np.array([s[2] for s in shapes])
    assert nc.min() >= 64 and nc.max() <= 2048 and side.min() >= 32 and side.max() <= 512
    assert set(s[1] for s in shapes) == {128, 256, 512}
