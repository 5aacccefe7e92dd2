"""GPU parity: the HIP correlate-sum path against the FP64 oracle.

Tolerances (relative L2 over the compared pixels, and the phase error at the
brightest oracle pixel):
  plain form      2e-3 / 2e-3 rad -- FP32 range and phase: the phase k*R of
                  ~1e3 cycles carries ~1e-4 cycles of rounding per term.
  precision form  1e-4 / 1e-4 rad -- the required bound; FP64 range/anchor.
  FP16 raw        3e-3            -- storage rounding 2^-11 per component.
  BF16 raw        2e-2            -- storage rounding 2^-8 per component.
  Doppler window  5e-3            -- terms within rounding of the band edge may
                  fall either side in FP32 vs FP64.
"""
import numpy as np
import pytest

from conftest import subset

pytestmark = pytest.mark.gpu

PLAIN_TOL = 1e-4  # rel-L2 bound of the metric, also for the drop-in form
PREC_TOL = 1e-4


def form(tdc, d=None, f64=None, center=0.0, bw=1e30, centroid=False, rw=False, taper=False,
         kind="plain", storage="c64", ppl=0, splits=0, engine="auto"):
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    if f64 is None:
        p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"], d["velocity"],
                      d["attitude"], d["modRate"], d["startFreq"], d["bodyBoresight"])
        p.setFocusGrid(d["focusGrid"])
    else:
        p.setRawDataF64(f64["rawData"], f64["priTimes"], f64["sampleTimes"], f64["position"],
                        f64["modRate"], f64["startFreq"])
        p.setFocusGridF64(f64["focusGrid"])
    p.setForm(kind)
    p.setRawStorage(storage)
    p.setLaunch(ppl, splits)
    p.setEngine(engine)
    p.start(center, bw, centroid, rw, taper)
    return p.getFocusedImage(), p.stats()


def oracle_dropin(orc, d, rw=None, doppler=None, pixels=None):
    return orc.image_f64(d["rawData"], d["sampleTimes"], d["position"], d["focusGrid"],
                         float(d["modRate"]), float(d["startFreq"]), range_win=rw,
                         doppler=doppler, pixels=pixels)


def oracle_f64(orc, s, pixels=None, rw=None):
    return orc.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate, s.f0,
                         range_win=rw, pixels=pixels)


def check(orc, img, ref, tol, ptol=None):
    e = orc.rel_l2(img, ref)
    ph = orc.peak_phase_error(img, ref)
    print(f"rel-L2 {e:.3e}  peak phase {ph:.3e} rad")
    assert e <= tol, e
    assert ph <= (ptol if ptol is not None else tol), ph
    return e, ph


def ragged_scene(npri=130, ns=100, rows=13, cols=21, seed=3):
    from sarlabtdc import scene

    rng = np.random.default_rng(seed)
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.2, 0.2, npri)
    track += rng.normal(scale=1e-4, size=track.shape)
    grid = scene.plane_grid(rows, cols, 4e-3, (0.01, 0.6, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 6)
    amp = scene._cn(rng, 6, 1.0)
    return scene.synthesize("ragged", track, ns, grid, scat, amp, 30.0, rng, speed=10.0)


def moving_within_chirp(d, s):
    """Per-sample positions: the antenna keeps moving along the track during each
    chirp (float32, as the reference API takes them)."""
    d = dict(d)
    vel = d["velocity"].astype(np.float64)
    t = s.sample_times[None, :, None]
    d["position"] = np.ascontiguousarray((d["position"].astype(np.float64) + vel * t)
                                         .astype(np.float32))
    return d


# ---------------------------------------------------------------------------- plain form


@pytest.mark.parametrize("motion,series", [(3e-4, True), (0.05, False)])
def test_plain_motion_within_chirps(tdc, orc, motion, series):
    """The antenna moves along x during every chirp (per-sample positions): by
    0.3 mm the range offset from the chirp's first position takes the two-term
    expansion, by 5 cm at 0.6 m the exact form num / (R_m + R0); both against
    the oracle at the metric's bound with the Doppler window off, and with it
    on at the windowed tolerance."""
    s = ragged_scene(npri=90, ns=128, rows=11, cols=13, seed=int(motion * 1e5))
    d = s.dropin()
    t = (s.sample_times - s.sample_times[0]) / (s.sample_times[-1] - s.sample_times[0])
    pos = d["position"].astype(np.float64)
    pos[:, :, 0] += motion * t[None, :]
    d["position"] = np.ascontiguousarray(pos.astype(np.float32))
    vel = np.zeros_like(d["velocity"])
    vel[..., 0] = 5.0
    d["velocity"] = vel
    img, st = form(tdc, d)
    assert not st["range_per_chirp"] and st["range_series"] == series
    check(orc, img, oracle_dropin(orc, d), PLAIN_TOL)
    w = dict(velocity=d["velocity"], attitude=d["attitude"], boresight=d["bodyBoresight"],
             center=0.0, bandwidth=2000.0, centroid=True, taper=True)
    img, st = form(tdc, d, center=0.0, bw=2000.0, centroid=True, taper=True)
    assert st["doppler_active"] and st["range_series"] == series
    # the window's band edge may fall between the FP32 and FP64 Doppler
    # frequencies of a term (as in the random drop-in tests)
    check(orc, img, oracle_dropin(orc, d, doppler=w), 5e-3)



def test_plain_line_small(tdc, orc):
    from sarlabtdc import scene

    s = scene.line_small()
    d = s.dropin()
    img, st = form(tdc, d)
    assert st["form"] == "plain" and not st["doppler_active"]
    assert img.shape == (32, 32) and img.dtype == np.complex64
    check(orc, img, oracle_dropin(orc, d), PLAIN_TOL)


def test_plain_line_small_hamming_range_window(tdc, orc):
    from sarlabtdc import scene

    d = scene.line_small().dropin()
    img, _ = form(tdc, d, rw=True)
    ref = oracle_dropin(orc, d, rw=orc.range_window(d["rawData"].shape[1]))
    check(orc, img, ref, PLAIN_TOL)


@pytest.mark.parametrize("window", [("hann", 0.0, 0.0), ("taylor", 4.0, 30.0)])
def test_plain_other_range_windows(tdc, orc, window):
    from sarlabtdc import scene

    d = scene.line_small().dropin()
    ns = d["rawData"].shape[1]
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"], d["velocity"],
                 d["attitude"], d["modRate"], d["startFreq"], d["bodyBoresight"])
    p.setFocusGrid(d["focusGrid"])
    p.setRangeWindow(*window)
    p.start(0.0, 1e30, False, True, False)
    img = p.getFocusedImage()
    if window[0] == "hann":
        rw = orc.range_window(ns, 0.5)
    else:
        rw = orc.taylor_window(ns, 4, 30.0)
    check(orc, img, oracle_dropin(orc, d, rw=rw), PLAIN_TOL)


@pytest.mark.parametrize("path", ["per_chirp", "per_term"])
@pytest.mark.parametrize("ppl", [1, 2, 4])
@pytest.mark.parametrize("splits", [0, 1, 3, 8])
def test_plain_launch_shapes_ragged(tdc, orc, ppl, splits, path):
    """Stop-and-hop positions take the per-(pixel, chirp) range; antenna motion
    within the chirps takes the per-term range."""
    s = ragged_scene()
    d = s.dropin()
    if path == "per_term":
        d = moving_within_chirp(d, s)
    img, st = form(tdc, d, ppl=ppl, splits=splits)
    assert st["pixels_per_lane"] == ppl
    assert st["range_per_chirp"] == (path == "per_chirp")
    assert not st["window_per_chirp"]
    if splits:
        assert st["chirp_splits"] == splits
    check(orc, img, oracle_dropin(orc, d), PLAIN_TOL)


@pytest.mark.parametrize("ppl", [1, 2, 4])
def test_plain_range_paths_bit_identical(tdc, ppl):
    """The per-chirp range is the per-term range computed once: the same image bit
    for bit. Sample 1 of every chirp carries zero raw data, so moving its position
    by one float32 ulp changes no term's contribution but sends the data down the
    per-term path."""
    s = ragged_scene()
    d = s.dropin()
    d["rawData"] = d["rawData"].copy()
    d["rawData"][:, 1] = 0
    a, sa = form(tdc, d, ppl=ppl, rw=True)
    d2 = dict(d)
    d2["position"] = d["position"].copy()
    d2["position"][:, 1, 0] = np.nextafter(d2["position"][:, 1, 0], np.float32(np.inf))
    b, sb = form(tdc, d2, ppl=ppl, rw=True)
    assert sa["range_per_chirp"] and not sb["range_per_chirp"]
    assert sa["chirp_splits"] == sb["chirp_splits"]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_plain_line_jitter_subset(tdc, orc):
    from sarlabtdc import scene

    s = scene.line_jitter()
    d = s.dropin()
    img, st = form(tdc, d)
    flat = img.reshape(-1)
    px = subset(flat.size, 768, extra=np.argsort(np.abs(flat))[-16:])
    ref = oracle_dropin(orc, d, pixels=px)
    check(orc, flat[px], ref, PLAIN_TOL)


def test_plain_point_target_closed_form(tdc):
    """One noiseless scatterer on a pixel, windows off: I(p0) = a * Nc * Ns."""
    from sarlabtdc import scene

    s = scene.line_small()
    p0 = s.grid[11, 17]
    amp = 0.8 - 0.6j
    raw = scene.simulate_beat(s.track, s.sample_times, p0[None], [amp]).astype(np.complex64)
    s.raw = raw
    d = s.dropin()
    d["focusGrid"] = np.ascontiguousarray(p0[None, None, :].astype(np.float32))
    # the oracle convention rounds the track to float32 too; use the float32 pixel
    img, _ = form(tdc, d)
    want = amp * raw.size
    assert abs(img[0, 0] - want) / abs(want) < 2e-3


def test_plain_linearity(tdc):
    from sarlabtdc import scene

    s = scene.line_small()
    d = s.dropin()
    rng = np.random.default_rng(11)
    r2 = (rng.normal(size=d["rawData"].shape) + 1j * rng.normal(size=d["rawData"].shape)).astype(
        np.complex64)
    i1, _ = form(tdc, d)
    d2 = dict(d, rawData=r2)
    i2, _ = form(tdc, d2)
    d3 = dict(d, rawData=(2.0 * d["rawData"] + r2).astype(np.complex64))
    i3, _ = form(tdc, d3)
    lhs = i3.astype(np.complex128)
    rhs = 2.0 * i1.astype(np.complex128) + i2
    assert np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs) < 1e-5


def test_plain_deterministic_reruns(tdc):
    from sarlabtdc import scene

    d = scene.line_small().dropin()
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"], d["velocity"],
                 d["attitude"], d["modRate"], d["startFreq"], d["bodyBoresight"])
    p.setFocusGrid(d["focusGrid"])
    p.start(0.0, 1e30, False, False, False)
    a = p.getFocusedImage()
    p.start(0.0, 1e30, False, False, False)
    b = p.getFocusedImage()
    p.prepare(0.0, 1e30, False, False, False)
    p.run(3)
    p.fetchImage()
    c = p.getFocusedImage()
    assert np.array_equal(a, b) and np.array_equal(a, c)


# ---------------------------------------------------------------------------- precision form


def test_precision_line_jitter_c64(tdc, orc):
    from sarlabtdc import scene

    s = scene.line_jitter()
    img, st = form(tdc, f64=s.f64(), kind="precision")
    assert st["form"] == "precision" and st["raw_storage"] == "c64"
    flat = img.reshape(-1)
    px = subset(flat.size, 768, extra=np.argsort(np.abs(flat))[-16:])
    check(orc, flat[px], oracle_f64(orc, s, pixels=px), PREC_TOL)


@pytest.mark.parametrize("storage,tol", [("f16", 3e-3), ("bf16", 2e-2)])
def test_precision_line_jitter_half_storage(tdc, orc, storage, tol):
    from sarlabtdc import scene

    s = scene.line_jitter()
    img, st = form(tdc, f64=s.f64(), kind="precision", storage=storage)
    assert st["raw_storage"] == storage
    flat = img.reshape(-1)
    px = subset(flat.size, 512, extra=np.argsort(np.abs(flat))[-16:])
    check(orc, flat[px], oracle_f64(orc, s, pixels=px), tol)


def test_precision_from_float_api(tdc, orc):
    """Precision form selected on the reference's float32 inputs (stop-and-hop)."""
    from sarlabtdc import scene

    d = scene.line_small().dropin()
    img, st = form(tdc, d, kind="precision")
    assert st["form"] == "precision"
    check(orc, img, oracle_dropin(orc, d), PREC_TOL)


@pytest.mark.parametrize("ppl", [1, 2, 4])
@pytest.mark.parametrize("splits", [0, 3])
def test_precision_ragged(tdc, orc, ppl, splits):
    """130 chirps, 100 samples (a 36-sample tail block), 13 x 21 pixels."""
    s = ragged_scene()
    img, _ = form(tdc, f64=s.f64(), kind="precision", ppl=ppl, splits=splits)
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("engine", ["vector", "matrix", "profile"])
@pytest.mark.parametrize("storage,tol", [("c64", PREC_TOL), ("f16", 3e-3), ("bf16", 2e-2)])
def test_precision_engines_ragged(tdc, orc, engine, storage, tol):
    s = ragged_scene()
    img, st = form(tdc, f64=s.f64(), kind="precision", storage=storage, engine=engine)
    assert st["engine"] == engine and st["raw_storage"] == storage
    check(orc, img, oracle_f64(orc, s), tol)


@pytest.mark.parametrize("ns", [1, 7, 64, 65, 127, 128, 129, 300, 512])
def test_matrix_engine_sample_counts(tdc, orc, ns):
    """Block-pair padding: 1..4 pairs, empty odd blocks, one-sample chirps."""
    s = ragged_scene(npri=40, ns=ns, rows=9, cols=11, seed=ns)
    img, st = form(tdc, f64=s.f64(), kind="precision", engine="matrix")
    assert st["engine"] == "matrix"
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("ns", [1, 2, 7, 8, 9, 64, 65, 127, 300, 512, 513, 1000, 1024, 1025,
                                1500, 2048, 2049, 3000, 4096, 4097, 6000, 8192])
def test_profile_engine_sample_counts(tdc, orc, ns):
    """FFT lengths 16..16384 (M >= 2 ns, at least 16), odd and even re-centring
    samples, one-sample chirps, Stockham FFTs whose two LDS buffers exceed 64
    KiB (M = 8192 from 2049 samples), the in-place FFT in one 128 KiB buffer
    (M = 16384 from 4097 samples), and the 8192-sample limit."""
    s = ragged_scene(npri=40, ns=ns, rows=9, cols=11, seed=ns)
    img, st = form(tdc, f64=s.f64(), kind="precision", engine="profile")
    assert st["engine"] == "profile"
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("spacing", [4e-3, 0.02, 0.15])
def test_profile_engine_window_and_gather_paths(tdc, orc, spacing):
    """Pixel spacings from 4 mm to 15 cm: a wave's taps fit the 64-bin window
    staged in LDS (close pixels), fit it for some chirps only, or never (a
    wave's 64 pixels spread over metres), where the taps are gathered from the
    table directly."""
    from sarlabtdc import scene

    rng = np.random.default_rng(31)
    npri, ns = 48, 256
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.3, 0.3, npri)
    track += rng.normal(scale=2e-4, size=track.shape)
    grid = scene.plane_grid(12, 70, spacing, (0.0, 1.2, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 6)
    s = scene.synthesize("spread", track, ns, grid, scat, scene._cn(rng, 6, 1.0), 30.0, rng)
    img, st = form(tdc, f64=s.f64(), kind="precision", engine="profile")
    assert st["engine"] == "profile"
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("taps,over,ns", [(8, 2, 300), (10, 2, 300), (12, 2, 300), (6, 4, 300),
                                          (8, 4, 300), (6, 4, 4096), (8, 4, 4096), (10, 2, 8192)])
def test_profile_engine_kernels(tdc, orc, taps, over, ns):
    """Every built interpolation kernel of the range-profile engine (taps,
    oversampling) against the oracle, each at the longest chirps its transform
    takes for the last two; the statistics name the kernel that ran."""
    s = ragged_scene(npri=40, ns=ns, rows=9, cols=11, seed=taps * over + ns)
    f = s.f64()
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"], f["modRate"],
                    f["startFreq"])
    p.setFocusGridF64(f["focusGrid"])
    p.setEngine("profile")
    p.setProfileKernel(taps, over)
    p.start(0.0, 1e30, False, False, False)
    st = p.stats()
    assert (st["engine"], st["taps"], st["oversampling"]) == ("profile", taps, over)
    check(orc, p.getFocusedImage(), oracle_f64(orc, s), PREC_TOL)


def test_profile_kernel_choice(tdc):
    """The automatic kernel: (6, 4) up to 4096 samples per chirp for images of
    at least 32768 pixels, (10, 2) for smaller images or longer chirps; a
    kernel that is not built, or fourfold oversampling beyond 4096 samples, is
    refused."""
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    with pytest.raises(ValueError):
        p.setProfileKernel(6, 2)
    with pytest.raises(ValueError):
        p.setProfileKernel(10, 4)
    for ns, side, want in ((4096, 182, (6, 4)), (4096, 181, (10, 2)), (4097, 182, (10, 2))):
        s = ragged_scene(npri=4, ns=ns, rows=side, cols=side, seed=ns + side)
        f = s.f64()
        p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"],
                        f["modRate"], f["startFreq"])
        p.setFocusGridF64(f["focusGrid"])
        p.setProfileKernel(0, 0)
        p.start(0.0, 1e30, False, False, False)
        st = p.stats()
        assert (st["taps"], st["oversampling"]) == want
    for taps in (6, 8):
        p.setProfileKernel(taps, 4)
        with pytest.raises(Exception, match="at most 4096 samples"):
            p.start(0.0, 1e30, False, False, False)


def test_profile_incremental_range(tdc, orc):
    """Within blocks of 8 chirps the range-profile tile advances each pixel's
    range, bin position and anchor phase by FP32 increments when the host's
    error bound allows it (config 2's 256 chirps, 2.7 mm steps at 0.8 m: it
    does; config 1's 64 chirps, 4.8 mm steps at 0.4 m: it does not), and every
    chirp's range in FP64 when asked. Both meet the oracle at 1e-4 and agree
    with each other to 2e-6."""
    from sarlabtdc import scene

    def run(s, exact):
        f = s.f64()
        p = tdc.TdcProcessor(0)
        p.setLogging(False)
        p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"],
                        f["modRate"], f["startFreq"])
        p.setFocusGridF64(f["focusGrid"])
        p.setProfileExactRange(exact)
        p.start(0.0, 1e30, False, False, False)
        return p.getFocusedImage().reshape(-1).copy(), p.stats()

    s = scene.line_jitter()
    inc, st_inc = run(s, False)
    ex, st_ex = run(s, True)
    assert st_inc["engine"] == "profile" and st_inc["incremental"]
    assert not st_ex["incremental"]
    agree = orc.rel_l2(inc, ex)
    px = subset(inc.size, 768, extra=np.argsort(np.abs(ex))[-16:])
    ref = oracle_f64(orc, s, pixels=px)
    print("incremental vs FP64 range", agree, "oracle", orc.rel_l2(inc[px], ref),
          orc.rel_l2(ex[px], ref))
    assert agree < 2e-6
    check(orc, inc[px], ref, PREC_TOL)
    check(orc, ex[px], ref, PREC_TOL)
    _, st_small = run(scene.line_small(), False)
    assert st_small["engine"] == "profile" and not st_small["incremental"]


def test_profile_engine_sample_limit(tdc, orc):
    """Beyond 8192 samples per chirp the range-profile engine declines with a
    message and the automatic choice takes the matrix engine."""
    s = ragged_scene(npri=12, ns=8193, rows=5, cols=6, seed=3)
    with pytest.raises(Exception, match="at most 8192 samples"):
        form(tdc, f64=s.f64(), kind="precision", engine="profile")
    img, st = form(tdc, f64=s.f64(), kind="precision")
    assert st["engine"] == "matrix"
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("engine", ["matrix", "profile"])
@pytest.mark.parametrize("gain_exp", [-30, 0, 30])
def test_matrix_engine_per_chirp_scale(tdc, orc, gain_exp, engine):
    """Chirp amplitudes spread over 1e-2..1e2 around 10**gain_exp: each chirp
    gets its own power-of-two FP16 scale."""
    s = ragged_scene(npri=61, ns=200, rows=7, cols=9, seed=11)
    rng = np.random.default_rng(1)
    gain = 10.0 ** (gain_exp + rng.uniform(-2, 2, s.n_chirps))
    s.raw = (s.raw * gain[:, None]).astype(np.complex64)
    img, _ = form(tdc, f64=s.f64(), kind="precision", engine=engine)
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("engine", ["vector", "matrix", "profile"])
def test_precision_down_chirp(tdc, orc, engine):
    """Negative chirp rate: the rotation step dk and the phase fractions are negative."""
    from sarlabtdc import scene

    rng = np.random.default_rng(21)
    npri, ns = 90, 300
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.2, 0.2, npri)
    grid = scene.plane_grid(11, 13, 4e-3, (0.0, 0.6, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 5)
    s = scene.synthesize("down", track, ns, grid, scat, scene._cn(rng, 5, 1.0), 30.0, rng,
                         f0=81e9, mod_rate=-scene.MOD_RATE)
    img, st = form(tdc, f64=s.f64(), kind="precision", engine=engine)
    assert st["engine"] == engine
    check(orc, img, oracle_f64(orc, s), PREC_TOL)


@pytest.mark.parametrize("engine", ["vector", "matrix", "profile"])
def test_precision_far_from_origin(tdc, orc, engine):
    """Coordinates 1e5 m from the origin (a projected map frame): FP64 ranges
    keep the phase; the image equals the one formed near the origin."""
    s = ragged_scene(npri=70, ns=200, rows=9, cols=12, seed=8)
    near, _ = form(tdc, f64=s.f64(), kind="precision", engine=engine)
    off = np.array([1.5e5, -2.3e5, 310.0])
    s.track = s.track + off
    s.grid = s.grid + off
    img, _ = form(tdc, f64=s.f64(), kind="precision", engine=engine)
    check(orc, img, oracle_f64(orc, s), PREC_TOL)
    assert orc.rel_l2(img, near) < 1e-5


@pytest.mark.parametrize("kind,engine", [("precision", "matrix"), ("precision", "vector"),
                                         ("precision", "profile"), ("plain", "auto")])
def test_more_splits_than_chirps(tdc, orc, kind, engine):
    """64 chirp ranges over 40 chirps: the empty ranges contribute zero partials."""
    s = ragged_scene(npri=40, ns=96, rows=5, cols=7, seed=17)
    if kind == "precision":
        img, st = form(tdc, f64=s.f64(), kind=kind, engine=engine, splits=64)
        tol = PREC_TOL
    else:
        img, st = form(tdc, s.dropin(), kind=kind, splits=64)
        tol = PLAIN_TOL
    assert st["chirp_splits"] == 64
    check(orc, img, oracle_f64(orc, s), tol)


def test_engines_agree(tdc, orc):
    from sarlabtdc import scene

    s = scene.line_jitter()
    a, _ = form(tdc, f64=s.f64(), kind="precision", engine="vector")
    b, _ = form(tdc, f64=s.f64(), kind="precision", engine="matrix")
    c, st = form(tdc, f64=s.f64(), kind="precision")
    assert st["engine"] == "profile"
    e = orc.rel_l2(b, a)
    e2 = orc.rel_l2(c, a)
    print(f"matrix vs vector rel-L2 {e:.3e}; profile vs vector {e2:.3e}")
    assert e < 1e-5 and e2 < 1e-5


def test_precision_range_window(tdc, orc):
    s = ragged_scene()
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    f = s.f64()
    p.setRawDataF64(f["rawData"], f["priTimes"], f["sampleTimes"], f["position"], f["modRate"],
                    f["startFreq"])
    p.setFocusGridF64(f["focusGrid"])
    p.start(0.0, 1e30, False, True, False)
    img = p.getFocusedImage()
    rw = orc.range_window(s.n_samples)
    check(orc, img, oracle_f64(orc, s, rw=rw), PREC_TOL)


def test_precision_point_target_closed_form(tdc):
    from sarlabtdc import scene

    s = scene.line_jitter(nc=128, ns=128, side=8)
    p0 = s.grid[3, 5]
    amp = -0.3 + 0.9j
    s.raw = scene.simulate_beat(s.track, s.sample_times, p0[None], [amp]).astype(np.complex64)
    s.grid = p0[None, None, :].copy()
    img, _ = form(tdc, f64=s.f64(), kind="precision")
    want = amp * s.raw.size
    assert abs(img[0, 0] - want) / abs(want) < 1e-5


def test_precision_rejects_nonuniform_times(tdc):
    s = ragged_scene()
    f = s.f64()
    f["sampleTimes"] = f["sampleTimes"] ** 1.01
    with pytest.raises(RuntimeError, match="uniformly spaced"):
        form(tdc, f64=f, kind="precision")


def test_precision_rejects_moving_phase_centre(tdc):
    d = ragged_scene().dropin()
    d["position"] = d["position"].copy()
    d["position"][:, :, 0] += np.linspace(0, 1e-4, d["position"].shape[1])[None, :]
    with pytest.raises(RuntimeError, match="stop-and-hop"):
        form(tdc, d, kind="precision")


# ---------------------------------------------------------------------------- Doppler window


def doppler_scene(npri=300, ns=64, side=16, speed=40.0, seed=5):
    from sarlabtdc import scene

    rng = np.random.default_rng(seed)
    track = np.zeros((npri, 3))
    track[:, 0] = np.linspace(-0.5, 0.5, npri)
    grid = scene.plane_grid(side, side, 5e-3, (0.0, 0.5, 0.0))
    scat = scene._uniform_on_grid(rng, grid, 5)
    amp = scene._cn(rng, 5, 1.0)
    s = scene.synthesize("doppler", track, ns, grid, scat, amp, 30.0, rng, speed=speed)
    return s


def doppler_dict(d, center, bw, centroid, taper):
    return dict(velocity=d["velocity"], attitude=d["attitude"], boresight=d["bodyBoresight"],
                center=center, bandwidth=bw, centroid=centroid, taper=taper)


@pytest.mark.parametrize("path", ["per_chirp", "per_term"])
@pytest.mark.parametrize("centroid,taper,center,bw", [
    (True, True, 0.0, 8e3),
    (False, False, 2e3, 6e3),
    (True, False, 0.0, 12e3),
])
def test_doppler_window_and_skip(tdc, orc, centroid, taper, center, bw, path):
    """Stop-and-hop data take the per-(pixel, chirp) window; antenna motion within
    the chirps takes the per-term window."""
    s = doppler_scene()
    d = s.dropin()
    if path == "per_term":
        d = moving_within_chirp(d, s)
    img, st = form(tdc, d, center=center, bw=bw, centroid=centroid, taper=taper)
    assert st["doppler_active"]
    assert st["window_per_chirp"] == (path == "per_chirp")
    assert st["range_per_chirp"] == (path == "per_chirp")
    dop = doppler_dict(d, center, bw, centroid, taper)
    ref = oracle_dropin(orc, d, doppler=dop)
    # skip counts against the oracle's own skip test
    g = d["focusGrid"].reshape(-1, 3).astype(np.float64)
    nchunks = (d["rawData"].shape[0] + 127) // 128
    processed = sum(
        orc.chunk_processed(d["rawData"], d["sampleTimes"], d["position"], float(d["modRate"]),
                            float(d["startFreq"]), dop, c, g[i])
        for i in range(g.shape[0]) for c in range(nchunks))
    assert st["processed_chunks"] + st["skipped_chunks"] == g.shape[0] * nchunks
    assert st["skipped_chunks"] > 0
    assert st["processed_chunks"] == processed
    check(orc, img, ref, 5e-3)


def test_doppler_matches_reference_method(tdc, orc):
    """GPU Doppler path vs the FP32 restatement of the reference method."""
    s = doppler_scene()
    d = s.dropin()
    img, _ = form(tdc, d, center=0.0, bw=8e3, centroid=True, taper=True, rw=True)
    rm = orc.refmethod_f32(d, 0.0, 8e3, True, True, True)
    assert orc.rel_l2(img, rm) < 5e-3


def test_doppler_window_paths_agree(tdc, orc):
    """The per-chirp and per-term windows on the same terms: one non-probe sample of
    every chirp moved by one float32 ulp sends the data down the per-term path."""
    s = doppler_scene()
    d = s.dropin()
    a, sa = form(tdc, d, center=0.0, bw=8e3, centroid=True, taper=True)
    d2 = dict(d)
    d2["position"] = d["position"].copy()
    d2["position"][:, 1, 0] = np.nextafter(d2["position"][:, 1, 0], np.float32(np.inf))
    b, sb = form(tdc, d2, center=0.0, bw=8e3, centroid=True, taper=True)
    assert sa["window_per_chirp"] and not sb["window_per_chirp"]
    assert sa["skipped_chunks"] == sb["skipped_chunks"] > 0
    e = orc.rel_l2(a, b)
    print(f"per-chirp vs per-term window: {e:.3e}")
    assert e < 1e-4


def test_doppler_trivial_window_uses_plain_kernel(tdc, orc):
    s = doppler_scene()
    d = s.dropin()
    img, st = form(tdc, d, center=0.0, bw=1e9, centroid=True, taper=False)
    assert not st["doppler_active"]
    check(orc, img, oracle_dropin(orc, d), PLAIN_TOL)


def test_pixel_on_antenna_position_takes_exact_path(tdc, orc):
    """A pixel exactly at an antenna position makes the reference's window 0/0."""
    s = doppler_scene(npri=140, side=4)
    d = s.dropin()
    g = d["focusGrid"].copy()
    g[0, 0] = d["position"][3, 0]
    d["focusGrid"] = g
    img, st = form(tdc, d, center=0.0, bw=1e9, centroid=True, taper=False)
    assert st["doppler_active"]
    dop = doppler_dict(d, 0.0, 1e9, True, False)
    ref = oracle_dropin(orc, d, doppler=dop)
    assert np.isfinite(img).all()
    assert orc.rel_l2(img, ref) < 5e-3


# ---------------------------------------------------------------------------- boundary behaviour


def test_empty_grid(tdc):
    d = ragged_scene().dropin()
    d["focusGrid"] = np.zeros((0, 5, 3), dtype=np.float32)
    img, _ = form(tdc, d)
    assert img.shape == (0, 5)


def test_single_term(tdc, orc):
    s = ragged_scene(npri=1, ns=1, rows=1, cols=1)
    d = s.dropin()
    img, _ = form(tdc, d)
    check(orc, img, oracle_dropin(orc, d), PLAIN_TOL)


def test_errors_mirror_reference(tdc):
    d = ragged_scene().dropin()
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    with pytest.raises(RuntimeError):
        p.getFocusedImage()
    with pytest.raises(RuntimeError, match="rawData must be 2D"):
        p.setRawData(d["rawData"][None], d["priTimes"], d["sampleTimes"], d["position"],
                     d["velocity"], d["attitude"], d["modRate"], d["startFreq"],
                     d["bodyBoresight"])
    with pytest.raises(RuntimeError, match="position shape is incorrect"):
        p.setRawData(d["rawData"], d["priTimes"], d["sampleTimes"], d["position"][:, :-1],
                     d["velocity"], d["attitude"], d["modRate"], d["startFreq"],
                     d["bodyBoresight"])
    with pytest.raises(RuntimeError, match="focusGrid must be 3D"):
        p.setFocusGrid(d["focusGrid"][0])
    with pytest.raises(ValueError, match="nPri must be greater than 0"):
        z = np.zeros((0, 4), np.complex64)
        p.setRawData(z, np.zeros(0, np.float32), np.zeros(4, np.float32),
                     np.zeros((0, 4, 3), np.float32), np.zeros((0, 4, 3), np.float32),
                     np.zeros((0, 4, 4), np.float32), 1.0, 1.0, d["bodyBoresight"])
    with pytest.raises(TypeError):
        p.setFocusGrid(d["focusGrid"].astype(np.float64))
    with pytest.raises(RuntimeError):
        tdc.TdcProcessor(9999)


def test_reference_smoke_all_zero(tdc):
    """The reference's SmokeTest (tdctest/tdctest.cpp:21-59): all-zero inputs."""
    n_rows = n_cols = 20
    npri, ns = 750, 1500
    p = tdc.TdcProcessor(0)
    p.setLogging(False)
    p.setFocusGrid(np.zeros((n_rows, n_cols, 3), np.float32))
    p.setRawData(np.zeros((npri, ns), np.complex64), np.zeros(npri, np.float32),
                 np.zeros(ns, np.float32), np.zeros((npri, ns, 3), np.float32),
                 np.zeros((npri, ns, 3), np.float32), np.zeros((npri, ns, 4), np.float32),
                 1.0, 1.0, np.array([0, 1, 0], np.float32))
    p.start(0.0, 5.0, True, True)
    img = p.getFocusedImage()
    st = p.stats()
    # pos = target = 0 makes every probe NaN: every (pixel, chunk) is skipped
    assert st["skipped_chunks"] == n_rows * n_cols * 6 and st["processed_chunks"] == 0
    assert np.array_equal(img, np.zeros((n_rows, n_cols), np.complex64))
