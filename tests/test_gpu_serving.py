"""Serving form: continuously batched jobs, parity, determinism in a mix, admission."""
import time

import numpy as np
import pytest

from conftest import subset

pytestmark = pytest.mark.gpu


def job_scene(j, nc, ns, side, spacing=3e-3):
    from sarlabtdc import scene

    return scene.stream_job(j, (nc, ns, side, spacing))


def submit(srv, s, form="precision", storage="c64"):
    f = s.f64()
    return srv.submit(f["rawData"], f["sampleTimes"], f["position"], f["focusGrid"],
                      f["modRate"], f["startFreq"], form=form, storage=storage)


def oracle_img(orc, s, px=None):
    return orc.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate, s.f0, pixels=px)


@pytest.mark.parametrize("form,storage,tol", [("precision", "c64", 1e-4), ("precision", "bf16", 2e-2),
                                              ("precision", "f16", 3e-3), ("plain", "c64", 1e-4)])
def test_single_job_parity(tdc, orc, form, storage, tol):
    s = job_scene(1, 200, 256, 70)
    srv = tdc.Server(0, 8e9)
    jid = submit(srv, s, form, storage)
    img, info = srv.wait(jid)
    srv.close()
    assert img.shape == (70, 70)
    px = subset(img.size, 600, extra=np.argsort(np.abs(img.reshape(-1)))[-8:])
    e = orc.rel_l2(img.reshape(-1)[px], oracle_img(orc, s, px))
    print(form, storage, "rel-L2", e, info)
    assert e < tol
    assert info["latency_ms"] > 0 and info["tiles"] >= 1


def test_long_chirp_job_parity(tdc, orc):
    """8200 samples per chirp: beyond the range-profile engine's 8192, the job
    runs on the matrix engine's tiles (their own persistent kernel)."""
    s = job_scene(2, 60, 8200, 40)
    srv = tdc.Server(0, 8e9)
    img, _ = srv.wait(submit(srv, s))
    srv.close()
    e = orc.rel_l2(img.reshape(-1), oracle_img(orc, s).reshape(-1))
    print("matrix-engine job rel-L2", e)
    assert e < 1e-4


def test_engines_bit_identical_alone_and_in_a_mix(tdc):
    """A range-profile job and a matrix-engine job, each alone and then in one
    mix with each other and a plain job: every image is the same bit for bit."""
    prof = job_scene(11, 150, 256, 80)
    mat = job_scene(12, 40, 8200, 48)
    plain = job_scene(13, 120, 128, 64)
    alone = []
    for s in (prof, mat):
        srv = tdc.Server(0, 8e9)
        alone.append(srv.wait(submit(srv, s))[0])
        srv.close()
    srv = tdc.Server(0, 8e9)
    ids = [submit(srv, mat), submit(srv, plain, form="plain"), submit(srv, prof)]
    mixed = [srv.wait(i)[0] for i in ids]
    srv.close()
    assert np.array_equal(alone[0], mixed[2])
    assert np.array_equal(alone[1], mixed[0])


def test_job_is_bit_identical_alone_and_in_a_mix(tdc):
    target = job_scene(7, 300, 256, 90)
    others = [job_scene(100 + k, nc, ns, side) for k, (nc, ns, side) in
              enumerate([(700, 512, 200), (64, 128, 32), (1500, 128, 64), (333, 256, 150),
                         (90, 512, 300)])]
    srv = tdc.Server(0, 16e9)
    alone, _ = srv.wait(submit(srv, target))
    srv.close()
    srv = tdc.Server(0, 16e9)
    ids = [submit(srv, o) for o in others[:3]]
    tid = submit(srv, target)
    ids += [submit(srv, o) for o in others[3:]]
    mixed, info = srv.wait(tid)
    for i in ids:
        srv.wait(i)
    st = srv.stats()
    srv.close()
    assert np.array_equal(alone, mixed)
    assert st["completed"] == 6 and st["failed"] == 0


@pytest.mark.parametrize("nc,ns,side,tiles", [(2048, 128, 16, 32), (512, 256, 64, 128),
                                              (600, 128, 256, 2560)])
def test_jobs_with_few_pixel_tiles_are_cut_into_short_chirp_ranges(tdc, orc, nc, ns, side, tiles):
    """A job with few 256-pixel tiles is cut into chirp ranges short enough
    for 8192 tiles, down to 64 chirps, while its partial images stay within
    four times its input bytes: 16^2 pixels x 2048 chirps gives one pixel
    tile x 32 ranges of 64 chirps, 64^2 x 512 gives 16 x 8, and 256^2 x 600
    (2.2 MB of input, 0.5 MB per FP32 partial image) 256 x 10 ranges of 64
    chirps.
    The image sums the partial images within the precision form's tolerance
    and is the same bit for bit alone and beside other jobs."""
    s = job_scene(21, nc, ns, side)
    srv = tdc.Server(0, 16e9)
    alone, info = srv.wait(submit(srv, s))
    srv.close()
    assert info["tiles"] == tiles, info
    px = subset(alone.size, 600, extra=np.argsort(np.abs(alone.reshape(-1)))[-8:])
    e = orc.rel_l2(alone.reshape(-1)[px], oracle_img(orc, s, px))
    print(nc, ns, side, "tiles", info["tiles"], "rel-L2", e)
    assert e < 1e-4
    srv = tdc.Server(0, 16e9)
    ids = [submit(srv, job_scene(200 + k, *shape)) for k, shape in
           enumerate([(700, 512, 200), (64, 128, 32)])]
    mixed, _ = srv.wait(submit(srv, s))
    for i in ids:
        srv.wait(i)
    srv.close()
    assert np.array_equal(alone, mixed)


def test_plain_job_bit_identical_in_a_mix(tdc):
    target = job_scene(8, 260, 128, 60)
    srv = tdc.Server(0, 8e9)
    alone, _ = srv.wait(submit(srv, target, form="plain"))
    srv.close()
    srv = tdc.Server(0, 8e9)
    a = submit(srv, job_scene(9, 800, 256, 128))
    b = submit(srv, target, form="plain")
    c = submit(srv, job_scene(10, 100, 512, 256), form="plain")
    mixed, _ = srv.wait(b)
    srv.wait(a)
    srv.wait(c)
    srv.close()
    assert np.array_equal(alone, mixed)


def test_admission_respects_the_budget(tdc, orc):
    jobs = [job_scene(20 + k, 256, 256, 128) for k in range(4)]
    srv = tdc.Server(0, 1.0)  # probe: how big is one job
    with pytest.raises(RuntimeError, match="more than the whole budget"):
        submit(srv, jobs[0])
    srv.close()
    one = None
    srv = tdc.Server(0, 64e9)
    img0, info = srv.wait(submit(srv, jobs[0]))
    one = info["device_bytes"]
    srv.close()
    # room for one job at a time: the others queue until memory frees up
    srv = tdc.Server(0, one * 1.5)
    ids = [submit(srv, s) for s in jobs]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["bytes_peak"] <= one * 1.5
    assert np.array_equal(res[0][0], img0)
    admits = [r[1]["t_admit"] for r in res]
    dones = [r[1]["t_done"] for r in res]
    for k in range(1, 4):
        assert admits[k] >= dones[k - 1] - 1e-3  # admitted only after the previous one left
    px = subset(128 * 128, 200)
    e = orc.rel_l2(res[3][0].reshape(-1)[px], oracle_img(orc, jobs[3], px))
    assert e < 1e-4


def test_few_pixel_jobs_under_a_tight_budget(tdc, orc):
    """Few-pixel jobs are cut into short chirp ranges, but their partial images
    stay within four times their input bytes: several such jobs (386 chirps x
    128 samples, 265^2 pixels) under a budget for two at a time all complete,
    within it, each the same image it forms alone."""
    jobs = [job_scene(60 + k, 386, 128, 265) for k in range(4)]
    srv = tdc.Server(0, 64e9)
    alone, info = srv.wait(submit(srv, jobs[0]))
    srv.close()
    one = info["device_bytes"]
    npx = 265 * 265
    inputs = 8 * 386 * 128 + 24 * npx
    nsplit = info["tiles"] / ((npx + 255) // 256)
    print("footprint", one, "splits", nsplit, "inputs", inputs)
    # FP32 partial images (the range-profile engine's)
    assert 8 * npx * nsplit <= 4 * inputs
    # arena: upload region, staged tables and partials, rounded up to a size class
    assert one <= 8 * inputs
    srv = tdc.Server(0, 2.5 * one)
    ids = [submit(srv, s) for s in jobs]
    res = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["bytes_peak"] <= 2.5 * one
    assert np.array_equal(res[0][0], alone)
    px = subset(npx, 200, extra=np.argsort(np.abs(res[2][0].reshape(-1)))[-4:])
    e = orc.rel_l2(res[2][0].reshape(-1)[px], oracle_img(orc, jobs[2], px))
    assert e < 1e-4


def test_arenas_are_recycled_within_the_budget(tdc):
    """Jobs of alternating sizes under a budget that holds one large job: the
    server reuses freed device blocks of a size and releases idle blocks of
    another size before allocating, never holding more than the budget, and
    every image equals the same job's image on a fresh server."""
    small = [job_scene(40 + k, 120, 128, 48) for k in range(3)]
    large = [job_scene(50 + k, 300, 256, 160) for k in range(3)]
    srv = tdc.Server(0, 64e9)
    alone = [srv.wait(submit(srv, s))[0] for s in (small[0], large[0])]
    big = srv.wait(submit(srv, large[1]))[1]["device_bytes"]
    srv.close()
    srv = tdc.Server(0, big * 1.2)
    order = [small[0], large[0], small[1], large[1], small[2], large[2]]
    res = []
    for s in order:  # one at a time: each size class is freed before the other returns
        res.append(srv.wait(submit(srv, s))[0])
    st = srv.stats()
    srv.close()
    assert st["completed"] == 6 and st["failed"] == 0
    assert st["bytes_peak"] <= big * 1.2
    assert np.array_equal(res[0], alone[0]) and np.array_equal(res[1], alone[1])


def test_small_jobs_overtake_a_large_one(tdc):
    """Jobs submitted while a large one runs join it on the device and finish first
    (the large job, 2048 chirps over 1536^2 pixels, takes ~10 ms of device time
    in several steps, so the small ones arrive well before its last step)."""
    big = job_scene(30, 2048, 512, 1536, spacing=1.5e-3)
    small = [job_scene(31 + k, 64, 128, 32) for k in range(5)]
    srv = tdc.Server(0, 64e9)
    # warm every upload thread, its stream and the pinned pools
    for i in [submit(srv, s) for s in small + small]:
        srv.wait(i)
    bid = submit(srv, big)
    time.sleep(0.003)
    sids = [submit(srv, s) for s in small]
    res_small = [srv.wait(i) for i in sids]
    _, binfo = srv.wait(bid)
    srv.close()
    for img, info in res_small:
        assert info["t_submit"] < binfo["t_done"]
        assert info["t_done"] < binfo["t_done"]
        assert info["t_first_step"] < binfo["t_done"]
    assert binfo["steps"] > 1


def test_invalid_jobs(tdc):
    s = job_scene(40, 64, 128, 32)
    f = s.f64()
    srv = tdc.Server(0, 8e9)
    with pytest.raises(ValueError, match="plain form reads complex64"):
        srv.submit(f["rawData"], f["sampleTimes"], f["position"], f["focusGrid"], f["modRate"],
                   f["startFreq"], form="plain", storage="bf16")
    bad_t = f["sampleTimes"] ** 1.01
    jid = srv.submit(f["rawData"], bad_t, f["position"], f["focusGrid"], f["modRate"],
                     f["startFreq"])
    with pytest.raises(RuntimeError, match="uniformly spaced"):
        srv.wait(jid)
    ok = submit(srv, s)
    img, _ = srv.wait(ok)
    srv.close()
    assert np.isfinite(img).all()
