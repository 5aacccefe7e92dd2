"""GPU parity through the C ABI itself (ctypes, no pybind): the calls a
cgo / JNI / ctypes binding of the reference would make, checked against the
golden fixtures and the FP64 oracle."""
import ctypes
import os

import numpy as np
import pytest

from test_abi import LIB, status_codes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

fp = ctypes.POINTER(ctypes.c_float)
dp = ctypes.POINTER(ctypes.c_double)


class Stats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("main_kernel_ms", ctypes.c_double),
                ("end_to_end_ms", ctypes.c_double), ("terms", ctypes.c_double),
                ("form", ctypes.c_int32), ("raw_storage", ctypes.c_int32),
                ("doppler_active", ctypes.c_int32), ("chirp_splits", ctypes.c_int32),
                ("pixels_per_lane", ctypes.c_int32), ("workgroups", ctypes.c_int32),
                ("skipped_chunks", ctypes.c_int64), ("processed_chunks", ctypes.c_int64),
                ("engine", ctypes.c_int32), ("per_chirp", ctypes.c_int32),
                ("profile_taps", ctypes.c_int32), ("profile_oversampling", ctypes.c_int32),
                ("profile_incremental", ctypes.c_int32), ("reserved", ctypes.c_int32)]


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.tdc_last_error.restype = ctypes.c_char_p
    lib.tdc_set_raw_data.argtypes = [ctypes.c_void_p, fp, fp, fp, fp, fp, fp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_float, ctypes.c_float, fp]
    lib.tdc_set_raw_data_f64.argtypes = [ctypes.c_void_p, fp, dp, dp, dp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.tdc_set_focus_grid.argtypes = [ctypes.c_void_p, fp, ctypes.c_int, ctypes.c_int]
    lib.tdc_set_focus_grid_f64.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int]
    lib.tdc_start.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                              ctypes.c_int, ctypes.c_int]
    lib.tdc_image_buffer.argtypes = [ctypes.c_void_p, ctypes.POINTER(fp)]
    lib.tdc_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
    assert lib.tdc_device_count() > 0, "no HIP device visible"
    return lib


def f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(fp)


def f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


def image(lib, h, rows, cols):
    ptr = fp()
    assert lib.tdc_image_buffer(h, ctypes.byref(ptr)) == 0, lib.tdc_last_error()
    flat = np.ctypeslib.as_array(ptr, shape=(rows * cols * 2,)).copy()
    return flat.view(np.complex64).reshape(rows, cols)


def test_capi_plain_golden_line_small(lib, orc):
    """setRawData / setFocusGrid / start / imageBuffer on the 64 x 128 fixture."""
    from golden.make_golden import load_dropin

    d, z = load_dropin("line_small_plain.npz")
    raw, p_raw = f32(np.ascontiguousarray(d["rawData"], dtype=np.complex64).view(np.float32))
    pri, p_pri = f32(d["priTimes"])
    t, p_t = f32(d["sampleTimes"])
    pos, p_pos = f32(d["position"])
    vel, p_vel = f32(d["velocity"])
    att, p_att = f32(d["attitude"])
    bore, p_b = f32(d["bodyBoresight"])
    grid, p_grid = f32(d["focusGrid"])
    nc, ns = d["rawData"].shape
    rows, cols = grid.shape[:2]
    h = ctypes.c_void_p()
    assert lib.tdc_create(0, ctypes.byref(h)) == 0, lib.tdc_last_error()
    try:
        assert lib.tdc_set_raw_data(h, p_raw, p_pri, p_t, p_pos, p_vel, p_att, nc, ns,
                                    float(d["modRate"]), float(d["startFreq"]), p_b) == 0
        assert lib.tdc_set_focus_grid(h, p_grid, rows, cols) == 0
        assert lib.tdc_start(h, 0.0, 1e30, 0, 0, 0) == 0, lib.tdc_last_error()
        img = image(lib, h, rows, cols)
        st = Stats()
        assert lib.tdc_get_stats(h, ctypes.byref(st)) == 0
    finally:
        lib.tdc_destroy(h)
    assert st.form == 0 and st.terms == rows * cols * nc * ns
    e = orc.rel_l2(img, z["image_f64"])
    print(f"plain via C ABI: rel-L2 {e:.3e}")
    assert e <= 2e-3


@pytest.mark.parametrize("engine", [1, 2, 3])
def test_capi_precision_golden_ragged(lib, orc, engine):
    """FP64 inputs, the three precision engines, on the 130 x 100 ragged fixture."""
    z = np.load(os.path.join(GOLDEN, "precision_ragged.npz"))
    codes = status_codes()
    raw, p_raw = f32(z["rawData"].view(np.float32))
    pri, p_pri = f64(z["priTimes"])
    t, p_t = f64(z["sampleTimes"])
    pos, p_pos = f64(z["position"])
    grid, p_grid = f64(z["focusGrid"])
    nc, ns = z["rawData"].shape
    rows, cols = grid.shape[:2]
    h = ctypes.c_void_p()
    assert lib.tdc_create(0, ctypes.byref(h)) == 0, lib.tdc_last_error()
    try:
        assert lib.tdc_set_raw_data_f64(h, p_raw, p_pri, p_t, p_pos, nc, ns,
                                        float(z["modRate"]), float(z["startFreq"])) == 0
        assert lib.tdc_set_focus_grid_f64(h, p_grid, rows, cols) == 0
        assert lib.tdc_set_engine(h, engine) == 0
        assert lib.tdc_set_engine(h, 7) == codes["TDC_ERR_INVALID_ARGUMENT"]
        assert lib.tdc_start(h, 0.0, 1e30, 0, 0, 0) == 0, lib.tdc_last_error()
        img = image(lib, h, rows, cols)
        st = Stats()
        assert lib.tdc_get_stats(h, ctypes.byref(st)) == 0
    finally:
        lib.tdc_destroy(h)
    assert st.form == codes["TDC_FORM_PRECISION"] and st.engine == engine
    e = orc.rel_l2(img, z["image_f64"])
    print(f"precision engine {engine} via C ABI: rel-L2 {e:.3e}")
    assert e <= 1e-4


def test_capi_rejects_sizes_beyond_int_range(lib):
    """Pixel counts above 2^30 and per-pixel term counts above 2^31 - 1 fail
    cleanly at the setter, before anything reads the (here too small) arrays."""
    codes = status_codes()
    bad = codes["TDC_ERR_INVALID_ARGUMENT"]
    buf, p_buf = f32(np.zeros(16))
    dbuf, p_dbuf = f64(np.zeros(16))
    h = ctypes.c_void_p()
    assert lib.tdc_create(0, ctypes.byref(h)) == 0, lib.tdc_last_error()
    try:
        assert lib.tdc_set_focus_grid(h, p_buf, 65536, 65536) == bad
        assert b"2^30" in lib.tdc_last_error()
        # one pixel beyond 2^30 (3 * index would pass 2^31 in 32-bit arithmetic)
        assert lib.tdc_set_focus_grid(h, p_buf, 32769, 32768) == bad
        assert lib.tdc_set_focus_grid_f64(h, p_dbuf, 46341, 46341) == bad
        assert lib.tdc_set_raw_data(h, p_buf, p_buf, p_buf, p_buf, p_buf, p_buf, 65536, 32768,
                                    1e14, 77e9, p_buf) == bad
        assert lib.tdc_set_raw_data_f64(h, p_buf, p_dbuf, p_dbuf, p_dbuf, 1 << 20, 4096,
                                        1e14, 77e9) == bad
        assert b"2^31" in lib.tdc_last_error()
        # at the limit the setters accept the sizes (nothing is read before a start)
        assert lib.tdc_set_focus_grid(h, p_buf, 32768, 32768) == 0
        assert lib.tdc_set_focus_grid_f64(h, p_dbuf, 32768, 32768) == 0
    finally:
        lib.tdc_destroy(h)
