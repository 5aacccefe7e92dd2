"""Seeded random geometries for the precision form: every engine against the
FP64 oracle, and tdc_start's streamed pass against the resident one.

Each case draws its own chirp count (1-200), samples per chirp (1-4096, with
sizes around powers of two), pixel grid (up to 1500 pixels, spacing 0.5 mm to
20 cm), stand-off range (0.2-30 m), aperture length, track jitter, up- or
down-chirp, and in some cases a scene shifted 2e5 m from the origin. The
tolerance is the precision form's 1e-4 rel-L2 for complex64 raw and the FP16
storage rounding (3e-3) for FP16 raw.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_CASES = 24
SAMPLE_CHOICES = [1, 2, 3, 5, 17, 64, 100, 128, 255, 256, 300, 511, 512, 700, 1000, 1024, 2049,
                  4096]


def random_scene(case):
    from sarlabtdc import scene

    rng = np.random.default_rng([2026, case])
    npri = int(rng.integers(1, 201))
    ns = int(rng.choice(SAMPLE_CHOICES))
    rows = int(rng.integers(1, 31))
    cols = int(rng.integers(1, min(90, 1500 // rows) + 1))
    spacing = float(np.exp(rng.uniform(np.log(5e-4), np.log(0.2))))
    stand_off = float(np.exp(rng.uniform(np.log(0.2), np.log(30.0))))
    track = np.zeros((npri, 3))
    length = float(rng.uniform(0.0, 2.0))
    track[:, 0] = np.linspace(-length / 2, length / 2, npri)
    track[:, 2] = rng.uniform(-0.3, 0.3)
    track += rng.normal(scale=float(rng.uniform(0, 2e-3)), size=track.shape)
    centre = (float(rng.uniform(-0.5, 0.5)), stand_off, float(rng.uniform(-1, 1)))
    grid = scene.plane_grid(rows, cols, spacing, centre)
    scat = scene._uniform_on_grid(rng, grid, 5)
    mod_rate = scene.MOD_RATE * (1 if rng.random() < 0.8 else -1)
    s = scene.synthesize("stress", track, ns, grid, scat, scene._cn(rng, 5, 1.0), 30.0, rng,
                         mod_rate=mod_rate)
    if rng.random() < 0.3:
        off = rng.uniform(-2e5, 2e5, size=3)
        s.track = s.track + off
        s.grid = s.grid + off
    desc = dict(npri=npri, ns=ns, rows=rows, cols=cols, spacing=spacing, range=stand_off,
                mod_rate=mod_rate)
    return s, desc


def processor(tdc, f, engine, storage):
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"], f["modRate"],
                    f["startFreq"])
    p.setFocusGridF64(f["focusGrid"])
    p.setEngine(engine)
    p.setRawStorage(storage)
    return p


@pytest.mark.parametrize("case", range(N_CASES))
def test_random_geometry(tdc, orc, case):
    s, desc = random_scene(case)
    print("case", case, desc, flush=True)
    f = s.f64()
    ref = orc.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate, s.f0).reshape(-1)
    engines = ["profile", "matrix", "vector"] if desc["ns"] <= 4096 else ["matrix", "vector"]
    storages = ["c64", "f16"] if case % 4 == 0 else ["c64"]
    for engine in engines:
        for storage in storages:
            p = processor(tdc, f, engine, storage)
            p.start(0.0, 1e30, False, False, False)
            img = p.getFocusedImage().copy()
            assert p.stats()["engine"] == engine
            e = float(orc.rel_l2(img.reshape(-1), ref))
            tol = 1e-4 if storage == "c64" else 3e-3
            print(f"  {engine}/{storage}: rel-L2 {e:.2e}", flush=True)
            assert e <= tol, (engine, storage, e)
            # the resident pass forms the same image bit for bit
            p.prepare(0.0, 1e30, False, False, False)
            p.run(1)
            p.fetchImage()
            again = p.getFocusedImage()
            assert np.array_equal(img.view(np.uint32), again.view(np.uint32))
