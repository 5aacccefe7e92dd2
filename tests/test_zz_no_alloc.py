"""The allocation-free entry point tdc_start_noalloc, through the C ABI
(ctypes): two calls in a row, the second with a smaller input and another raw
storage, each bit-identical to tdc_start's image with the same chirp split
count (same kernels, same host tables), the first also within 1e-4 of the
FP64 oracle; inputs too large for the static scratch array and a too small
output buffer are refused with a message."""
import ctypes

import numpy as np
import pytest

from conftest import subset
from test_abi import LIB, status_codes
from test_gpu_capi import Stats, dp, f32, f64, fp, image

pytestmark = pytest.mark.gpu

RAW = {"c64": 0, "f16": 1, "bf16": 2}


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.tdc_last_error.restype = ctypes.c_char_p
    lib.tdc_set_raw_data_f64.argtypes = [ctypes.c_void_p, fp, dp, dp, dp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.tdc_set_focus_grid_f64.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int]
    lib.tdc_start.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
    lib.tdc_start_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, fp,
                                      ctypes.c_size_t]
    lib.tdc_image_buffer.argtypes = [ctypes.c_void_p, ctypes.POINTER(fp)]
    lib.tdc_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
    lib.tdc_set_launch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.tdc_set_raw_storage.argtypes = [ctypes.c_void_p, ctypes.c_int]
    assert lib.tdc_device_count() > 0, "no HIP device visible"
    return lib


WIN_OFF = (0.0, 1e30, 0, 0, 0)


class Inputs:
    """FP64 inputs of a scene, kept alive while the processor borrows them."""

    def __init__(self, s):
        f = s.f64()
        self.raw, self.p_raw = f32(np.ascontiguousarray(f["rawData"]).view(np.float32))
        self.t, self.p_t = f64(f["sampleTimes"])
        self.pri, self.p_pri = f64(f["priTimes"])
        self.pos, self.p_pos = f64(f["position"])
        self.grid, self.p_grid = f64(f["focusGrid"])
        self.nc, self.ns = f["rawData"].shape
        self.rows, self.cols = f["focusGrid"].shape[:2]
        self.mod_rate, self.f0 = f["modRate"], f["startFreq"]

    def bind(self, lib, h, storage):
        assert lib.tdc_set_raw_data_f64(h, self.p_raw, self.p_pri, self.p_t, self.p_pos, self.nc,
                                        self.ns, self.mod_rate, self.f0) == 0, lib.tdc_last_error()
        assert lib.tdc_set_focus_grid_f64(h, self.p_grid, self.rows, self.cols) == 0
        assert lib.tdc_set_raw_storage(h, RAW[storage]) == 0


def noalloc(lib, h, x):
    out = np.zeros(2 * x.rows * x.cols, dtype=np.float32)
    rc = lib.tdc_start_noalloc(h, *WIN_OFF, out.ctypes.data_as(fp), out.size)
    assert rc == 0, lib.tdc_last_error()
    st = Stats()
    assert lib.tdc_get_stats(h, ctypes.byref(st)) == 0
    assert st.engine == 3 and st.form == 1  # range-profile engine, precision form
    return out.view(np.complex64).reshape(x.rows, x.cols), st.chirp_splits


def start(lib, h, x, storage, splits):
    x.bind(lib, h, storage)
    assert lib.tdc_set_launch(h, 0, splits) == 0
    assert lib.tdc_start(h, *WIN_OFF) == 0, lib.tdc_last_error()
    st = Stats()
    assert lib.tdc_get_stats(h, ctypes.byref(st)) == 0
    assert st.chirp_splits == splits and st.engine == 3
    return image(lib, h, x.rows, x.cols)


def test_noalloc_matches_start_twice(lib, orc):
    from sarlabtdc import scene

    big = Inputs(scene.arc())              # 1024 chirps x 512 samples, 512^2 pixels
    small = Inputs(scene.line_jitter())    # 256 chirps x 256 samples, 128^2 pixels
    h = ctypes.c_void_p()
    ref = ctypes.c_void_p()
    assert lib.tdc_create(0, ctypes.byref(h)) == 0, lib.tdc_last_error()
    assert lib.tdc_create(0, ctypes.byref(ref)) == 0, lib.tdc_last_error()
    try:
        print("noalloc: arc c64, then line_jitter bf16", flush=True)
        big.bind(lib, h, "c64")
        img_big, s_big = noalloc(lib, h, big)
        small.bind(lib, h, "bf16")
        img_small, s_small = noalloc(lib, h, small)
        print(f"splits {s_big} (arc), {s_small} (line_jitter)", flush=True)
        assert s_big >= 1 and s_small >= 1
        want_big = start(lib, ref, big, "c64", s_big)
        want_small = start(lib, ref, small, "bf16", s_small)
        assert np.array_equal(img_big.view(np.uint32), want_big.view(np.uint32))
        assert np.array_equal(img_small.view(np.uint32), want_small.view(np.uint32))
        # and the image itself against the FP64 oracle
        flat = img_big.reshape(-1)
        px = subset(flat.size, 256, seed=5, extra=np.argsort(np.abs(flat))[-8:])
        s = scene.arc()
        o = orc.image_f64(s.raw, s.sample_times, s.track, s.grid, s.mod_rate, s.f0, pixels=px)
        e = orc.rel_l2(flat[px], o)
        print(f"arc noalloc rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-4
        # a third call on the processor after tdc_start: still the same image
        big.bind(lib, h, "c64")
        again, s_again = noalloc(lib, h, big)
        assert s_again == s_big
        assert np.array_equal(again.view(np.uint32), want_big.view(np.uint32))
    finally:
        lib.tdc_destroy(h)
        lib.tdc_destroy(ref)


def test_noalloc_refuses_what_does_not_fit(lib):
    bad = status_codes()["TDC_ERR_INVALID_ARGUMENT"]
    nc, ns, side = 8192, 512, 16
    raw, p_raw = f32(np.zeros((nc, ns, 2)))
    t, p_t = f64(np.arange(ns) * 40e-6 / ns)
    pri, p_pri = f64(np.arange(nc) * 1e-4)
    pos, p_pos = f64(np.stack([np.linspace(-0.5, 0.5, nc), np.zeros(nc), np.zeros(nc)], 1))
    g = np.zeros((side, side, 3))
    g[..., 1] = 2.0
    grid, p_grid = f64(g)
    h = ctypes.c_void_p()
    assert lib.tdc_create(0, ctypes.byref(h)) == 0, lib.tdc_last_error()
    try:
        assert lib.tdc_set_raw_data_f64(h, p_raw, p_pri, p_t, p_pos, nc, ns, 1e14, 77e9) == 0
        assert lib.tdc_set_focus_grid_f64(h, p_grid, side, side) == 0
        out = np.zeros(2 * side * side, dtype=np.float32)
        # 8192 chirps of 512 samples: 33.5 MB of raw and 68 MB of tables
        rc = lib.tdc_start_noalloc(h, *WIN_OFF, out.ctypes.data_as(fp), out.size)
        assert rc == bad
        assert b"scratch" in lib.tdc_last_error()
        # an output buffer smaller than the image
        assert lib.tdc_set_raw_data_f64(h, p_raw, p_pri, p_t, p_pos, 64, ns, 1e14, 77e9) == 0
        rc = lib.tdc_start_noalloc(h, *WIN_OFF, out.ctypes.data_as(fp), out.size - 1)
        assert rc == bad
        assert b"floats" in lib.tdc_last_error()
        # the same inputs with the full buffer form an image (all zeros here)
        rc = lib.tdc_start_noalloc(h, *WIN_OFF, out.ctypes.data_as(fp), out.size)
        assert rc == 0, lib.tdc_last_error()
        assert not np.any(out)
    finally:
        lib.tdc_destroy(h)
