"""tdc_start's streamed ingest against the resident path.

start() uploads the chirp inputs first and the focus grid behind them in pixel
slabs, launching each slab's correlate-sum as soon as its part of the grid has
landed. prepare() + run(1) + fetchImage() uploads everything and forms the
image in one launch. Every pixel goes through the same tile code over the same
chirp splits either way, so the two images must agree bit for bit, and the
Doppler skip counters must agree exactly.

The scene is sized so that start() really streams: 512 x 512 pixels (1024
tiles of 256) over 256 chirps gives 16 chirp splits, and the grid is cut in two
slabs once tiles x splits reaches 8192.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WIN_OFF = (0.0, 1e30, False, False, False)
# a narrow centroid-centred band with the taper: the window varies over the data
WIN_DOPPLER = (0.0, 150.0, True, False, True)


@pytest.fixture(scope="module")
def jitter():
    from sarlabtdc import scene

    return scene.make("line_jitter", side=512)


def per_term_positions(d, s):
    """The antenna keeps moving along its velocity during each chirp."""
    d = dict(d)
    vel = d["velocity"].astype(np.float64)
    t = s.sample_times[None, :, None]
    d["position"] = np.ascontiguousarray((d["position"].astype(np.float64) + vel * t)
                                         .astype(np.float32))
    return d


def processor(tdc, d=None, f=None, kind="plain", storage="c64", engine="auto"):
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    if f is None:
        p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"], d["velocity"],
                     d["attitude"], d["modRate"], d["startFreq"], d["bodyBoresight"])
        p.setFocusGrid(d["focusGrid"])
    else:
        p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"],
                        f["modRate"], f["startFreq"])
        p.setFocusGridF64(f["focusGrid"])
    p.setForm(kind)
    p.setRawStorage(storage)
    p.setEngine(engine)
    return p


def resident(p, win):
    p.prepare(*win)
    p.run(1)
    p.fetchImage()
    return p.getFocusedImage().copy(), p.stats()


def streamed(p, win):
    p.start(*win)
    return p.getFocusedImage().copy(), p.stats()


def assert_same(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
        f"{np.count_nonzero(a != b)} pixels differ"


def assert_streams(st, npx=512 * 512):
    tiles = -(-npx // (256 * st["pixels_per_lane"]))
    assert tiles * st["chirp_splits"] >= 8192, "scene too small to stream the grid"


CASES = [
    # (id, input, form, storage, engine, window)
    ("plain_per_chirp", "dropin", "plain", "c64", "auto", WIN_OFF),
    ("plain_per_term", "moving", "plain", "c64", "auto", WIN_OFF),
    ("doppler_per_chirp", "dropin", "plain", "c64", "auto", WIN_DOPPLER),
    ("doppler_per_term", "moving", "plain", "c64", "auto", WIN_DOPPLER),
    ("profile_c64", "f64", "precision", "c64", "profile", WIN_OFF),
    ("profile_f16", "f64", "precision", "f16", "profile", WIN_OFF),
    ("matrix_c64", "f64", "precision", "c64", "matrix", WIN_OFF),
    ("vector_c64", "f64", "precision", "c64", "vector", WIN_OFF),
    ("profile_fp32_api", "dropin", "precision", "c64", "auto", WIN_OFF),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_start_streams_the_same_image_as_prepare_run(tdc, jitter, case):
    name, src, kind, storage, engine, win = case
    s = jitter
    if src == "f64":
        p = processor(tdc, f=s.f64(), kind=kind, storage=storage, engine=engine)
    else:
        d = s.dropin() if src == "dropin" else per_term_positions(s.dropin(), s)
        p = processor(tdc, d=d, kind=kind, storage=storage, engine=engine)
    a, sa = streamed(p, win)
    b, sb = resident(p, win)
    print(name, sa["form"], sa["engine"], sa["chirp_splits"], sa["doppler_active"],
          sa["range_per_chirp"], sa["skipped_chunks"])
    assert_streams(sa)
    assert sa["form"] == kind
    if win is WIN_DOPPLER:
        assert sa["doppler_active"]
        assert sa["range_per_chirp"] == (src == "dropin")
        assert (sa["processed_chunks"], sa["skipped_chunks"]) == \
            (sb["processed_chunks"], sb["skipped_chunks"])
    assert_same(a, b)
    # and start() again on the same processor (the ingest's slots are reused)
    c, _ = streamed(p, win)
    assert_same(a, c)


def test_start_after_new_inputs_matches_a_fresh_processor(tdc, jitter):
    """Two start() calls with different inputs on one processor: the second
    image is the one a fresh processor forms from those inputs."""
    from sarlabtdc import scene

    first = jitter.dropin()
    other = scene.make("line_jitter", seed=5, side=512).dropin()
    p = processor(tdc, d=first)
    streamed(p, WIN_OFF)
    p.setRawData(other["rawData"], other["priTimes"], other["sampleTimes"], other["position"],
                 other["velocity"], other["attitude"], other["modRate"], other["startFreq"],
                 other["bodyBoresight"])
    p.setFocusGrid(other["focusGrid"])
    a, _ = streamed(p, WIN_OFF)
    b, _ = resident(processor(tdc, d=other), WIN_OFF)
    assert_same(a, b)


def test_small_image_runs_unstreamed(tdc):
    """Below the streaming threshold start() forms the image in one launch; the
    result is the same as prepare + run."""
    from sarlabtdc import scene

    d = scene.line_small().dropin()
    p = processor(tdc, d=d)
    a, _ = streamed(p, WIN_OFF)
    b, _ = resident(p, WIN_OFF)
    assert_same(a, b)


@pytest.mark.parametrize("kind,ppl,splits", [("plain", 2, 24), ("plain", 4, 40), ("plain", 1, 12),
                                             ("precision", 2, 20), ("precision", 0, 12)])
def test_start_streams_forced_launch_shapes(tdc, jitter, kind, ppl, splits):
    """Forced pixels per lane (tiles of 512 or 1024 pixels) and chirp split
    counts that are not powers of two (no XCD-aware mapping): the grid slabs
    are cut at whole tiles and the partial images stay one full image apart,
    so start() still forms the resident pass's image bit for bit."""
    s = jitter
    if kind == "precision":
        p = processor(tdc, f=s.f64(), kind=kind, engine="vector" if ppl > 1 else "auto")
    else:
        p = processor(tdc, d=s.dropin(), kind=kind)
    p.setLaunch(ppl, splits)
    a, sa = streamed(p, WIN_OFF)
    b, _ = resident(p, WIN_OFF)
    assert sa["chirp_splits"] == splits
    assert_streams(sa)
    assert_same(a, b)


# ---------------------------------------------------------------- device budget

@pytest.fixture(scope="module")
def long_aperture():
    """2048 chirps x 512 samples on a 64 x 64 grid: ~34 MB of per-chirp data
    (raw, staged raw, range-profile tables) against ~1 MB per pixel buffer."""
    from sarlabtdc import scene

    rng = np.random.default_rng(77)
    npri, ns = 2048, 512
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.4, 0.4, npri)
    track += rng.normal(scale=2e-4, size=track.shape)
    grid = scene.plane_grid(64, 64, 3e-3, (0.0, 1.0, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 40)
    return scene.synthesize("long", track, ns, grid, scat, scene._cn(rng, 40, 1.0), 25.0, rng)


BUDGET = 12 << 20


@pytest.mark.parametrize("engine,storage", [("profile", "c64"), ("profile", "f16"),
                                            ("matrix", "c64"), ("vector", "c64")])
def test_budget_streams_chirp_slabs(tdc, orc, long_aperture, engine, storage):
    """An aperture larger than the device budget: start() streams chirp slabs
    through the device and forms, bit for bit, the image of a resident pass
    with the same chirp splits; the image passes the oracle."""
    s = long_aperture
    f = s.f64()
    p = processor(tdc, f=f, kind="precision", storage=storage, engine=engine)
    p.setDeviceBudget(BUDGET)
    a, sa = streamed(p, WIN_OFF)
    nsplit = sa["chirp_splits"]
    print(engine, storage, "splits", nsplit, "engine", sa["engine"])
    assert sa["engine"] == engine
    # the per-chirp data alone exceed the budget: it cannot have been resident
    per_chirp = 8 * 512 + (8 if storage == "c64" else 4) * 512
    assert per_chirp * 2048 > BUDGET or engine != "vector"
    # prepare() keeps an aperture resident and refuses this one
    with pytest.raises(Exception, match="device budget"):
        p.prepare(*WIN_OFF)
    q = processor(tdc, f=f, kind="precision", storage=storage, engine=engine)
    q.setLaunch(0, nsplit)
    b, sb = resident(q, WIN_OFF)
    assert sb["chirp_splits"] == nsplit
    assert_same(a, b)
    px = np.unique(np.random.default_rng(5).choice(64 * 64, 300, replace=False))
    ref = orc.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate, s.f0, pixels=px)
    e = orc.rel_l2(a.reshape(-1)[px], ref)
    assert e <= (1e-4 if storage == "c64" else 3e-3), e


def test_budget_too_small_and_plain_form(tdc, long_aperture):
    """A budget below the per-pixel buffers (grid, image, FP64 sum and one
    FP32 partial image: 56 B a pixel, 224 KiB here) holds no slab and fails with a
    message; 1 MiB already streams slabs of 8 chirps. The plain form keeps its
    aperture resident whatever the budget."""
    s = long_aperture
    p = processor(tdc, f=s.f64(), kind="precision")
    p.setDeviceBudget(200 << 10)
    with pytest.raises(Exception, match="holds no slab"):
        p.start(*WIN_OFF)
    p.setDeviceBudget(1 << 20)
    a, sa = streamed(p, WIN_OFF)
    q = processor(tdc, f=s.f64(), kind="precision")
    q.setLaunch(0, sa["chirp_splits"])
    b, _ = resident(q, WIN_OFF)
    assert_same(a, b)
    d = s.dropin()
    q = processor(tdc, d=d, kind="plain")
    q.setDeviceBudget(1 << 20)
    a, _ = streamed(q, WIN_OFF)
    r = processor(tdc, d=d, kind="plain")
    b, _ = resident(r, WIN_OFF)
    assert_same(a, b)


def test_budget_streams_fp32_api(tdc, long_aperture):
    """The reference's FP32 inputs in the precision form under a budget: the
    chirp slabs form the resident pass's image bit for bit."""
    d = long_aperture.dropin()
    p = processor(tdc, d=d, kind="precision")
    p.setDeviceBudget(BUDGET)
    a, sa = streamed(p, WIN_OFF)
    assert sa["form"] == "precision" and sa["engine"] == "profile"
    q = processor(tdc, d=d, kind="precision")
    q.setLaunch(0, sa["chirp_splits"])
    b, _ = resident(q, WIN_OFF)
    assert_same(a, b)
