"""The C ABI library: it loads, exports every symbol its header declares, and
reports errors through status codes and tdc_last_error() (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sarlabtdc_amd.h")
LIB = os.path.join(ROOT, "sarlab-tdc_amd", "sarlabtdc", "libsarlabtdc_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tdc_[a-z0-9_]+)\s*\(", text)))


def status_codes():
    text = open(HEADER).read()
    return {k: int(v) for k, v in re.findall(r"#define (TDC_[A-Z0-9_]+) (\d+)", text)}


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.tdc_last_error.restype = ctypes.c_char_p
    return lib


def test_header_declares_the_reference_boundary():
    names = declared_functions()
    for n in ("tdc_create", "tdc_destroy", "tdc_set_raw_data", "tdc_set_focus_grid",
              "tdc_start", "tdc_image_buffer", "tdc_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_device_count(lib):
    # 2.0.0: tdc_stats grew by the range-profile kernel fields
    assert lib.tdc_abi_version() == 20000
    assert lib.tdc_device_count() >= 0


def test_null_handle_is_invalid_argument(lib):
    codes = status_codes()
    bad = codes["TDC_ERR_INVALID_ARGUMENT"]
    assert lib.tdc_set_focus_grid(None, None, 0, 0) == bad
    assert b"NULL" in lib.tdc_last_error()
    assert lib.tdc_start(None, ctypes.c_float(0), ctypes.c_float(1), 0, 0, 0) == bad
    assert lib.tdc_run(None, 1) == bad
    assert lib.tdc_start_noalloc(None, ctypes.c_float(0), ctypes.c_float(1), 0, 0, 0, None,
                                 ctypes.c_size_t(0)) == bad
    assert lib.tdc_destroy(None) == codes["TDC_OK"]
    assert lib.tdc_last_error() == b""


def test_create_without_device_reports_device_error(lib):
    if lib.tdc_device_count() > 0:
        pytest.skip("a HIP device is visible")
    h = ctypes.c_void_p()
    assert lib.tdc_create(0, ctypes.byref(h)) == status_codes()["TDC_ERR_DEVICE"]
    assert h.value is None
    assert b"Failed to set HIP device" in lib.tdc_last_error()


def test_python_module_without_device_raises_runtime_error():
    import sarlabtdc

    if sarlabtdc.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(RuntimeError, match="Failed to set HIP device"):
        sarlabtdc.TdcProcessor(0)


def test_python_surface_matches_reference():
    """Method and keyword names of py-sarlabtdc/src/interface.cpp:15-22."""
    import sarlabtdc

    cls = sarlabtdc.TdcProcessor
    for name in ("start", "setRawData", "setFocusGrid", "getFocusedImage"):
        assert hasattr(cls, name)
    doc = cls.start.__doc__
    for kw in ("dopplerWinCenter", "dopplerBandwidth", "dopCentroidWin", "applyRangeWin",
               "dopplerWinTaper: bool = True"):
        assert kw in doc


def test_cpp_header_compiles_against_library(tmp_path):
    """A C++ user of tdcprocessor.h (the reference's tdctest SmokeTest shape) builds."""
    src = tmp_path / "user.cpp"
    src.write_text(r'''
#include <complex>
#include <vector>
#include "tdcprocessor.h"
int main() {
    TdcProcessor tdc(0);
    std::vector<float> grid(20 * 20 * 3);
    tdc.setFocusGrid(grid.data(), 20, 20);
    const int nPri = 750, nSamples = 1500;
    std::vector<std::complex<float>> raw(nPri * nSamples);
    std::vector<float> priTimes(nPri), sampleTimes(nSamples), position(raw.size() * 3),
        velocity(raw.size() * 3), attitude(raw.size() * 4);
    float3 bore; bore.x = 0; bore.y = 1; bore.z = 0;
    tdc.setRawData(raw.data(), priTimes.data(), sampleTimes.data(), position.data(),
                   velocity.data(), attitude.data(), nPri, nSamples, 1.0f, 1.0f, bore);
    tdc.setLoggerSinks({std::make_shared<tdclog::stdout_sink>()});
    tdc.start(0.0f, 5.0f, true, true);
    return tdc.imageBuffer() == nullptr;
}
''')
    import subprocess

    pkg = os.path.join(ROOT, "sarlab-tdc_amd", "sarlabtdc")
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I/opt/rocm/include", str(src), "-L", pkg,
           "-lsarlabtdc_amd", "-Wl,-rpath," + pkg, "-o", str(tmp_path / "user")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("ns", [1, 7, 64, 100, 512])
def test_range_window_matches_scipy(lib, ns):
    """The library's FP64 range windows against scipy, on the host. Periodic =
    the symmetric (ns+1)-point window without its last point (scipy's sym=False,
    written out so that ns = 1 gives the reference's A - (1-A) cos 0, not 1)."""
    import numpy as np
    from scipy.signal import windows

    def periodic(f, *a, **k):
        return f(ns + 1, *a, sym=True, **k)[:-1]

    lib.tdc_range_window.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]
    codes = status_codes()
    cases = [
        (codes["TDC_WIN_HAMMING"], 25.0 / 46.0, 0.0,
         periodic(windows.general_hamming, 25.0 / 46.0)),
        (codes["TDC_WIN_HANN"], 0.0, 0.0, periodic(windows.hann)),
        (codes["TDC_WIN_TAYLOR"], 4.0, 30.0, periodic(windows.taylor, nbar=4, sll=30)),
        (codes["TDC_WIN_TAYLOR"], 6.0, 40.0, periodic(windows.taylor, nbar=6, sll=40)),
    ]
    for kind, p0, p1, want in cases:
        out = np.empty(ns, np.float32)
        rc = lib.tdc_range_window(kind, p0, p1, ns,
                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert rc == codes["TDC_OK"], lib.tdc_last_error()
        np.testing.assert_allclose(out, want.astype(np.float32), rtol=0, atol=2e-7)
    bad = np.empty(ns, np.float32)
    p = bad.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.tdc_range_window(99, 0.0, 0.0, ns, p) == codes["TDC_ERR_INVALID_ARGUMENT"]
    assert lib.tdc_range_window(codes["TDC_WIN_TAYLOR"], 0.5, 30.0, ns, p) == \
        codes["TDC_ERR_INVALID_ARGUMENT"]
    assert lib.tdc_range_window(codes["TDC_WIN_HANN"], 0.0, 0.0, 0, p) == \
        codes["TDC_ERR_INVALID_ARGUMENT"]


def test_cpp_header_takes_spdlog_sinks(tmp_path):
    """When the including translation unit sees <spdlog/spdlog.h>, the header
    declares the reference's setLoggerSinks(std::vector<spdlog::sink_ptr>)
    (tdc-lib/include/tdcprocessor.h:128) and forwards each line to the spdlog
    sink. spdlog is not installed here, so the test compiles against a minimal
    header with spdlog's sink interface (sinks::sink::log(const
    details::log_msg &), level_enum, sink_ptr) and checks the forwarding on the
    host, without a device."""
    inc = tmp_path / "inc" / "spdlog"
    inc.mkdir(parents=True)
    (inc / "spdlog.h").write_text(r"""
#pragma once
#include <memory>
#include <string>
namespace spdlog {
namespace level { enum level_enum { trace, debug, info, warn, err, critical, off }; }
namespace details {
struct log_msg {
    log_msg(const std::string &n, level::level_enum l, const std::string &m)
        : logger_name(n), level(l), payload(m) {}
    std::string logger_name;
    level::level_enum level;
    std::string payload;
};
}
namespace sinks {
class sink {
public:
    virtual ~sink() = default;
    virtual void log(const details::log_msg &msg) = 0;
    virtual void flush() = 0;
};
}
using sink_ptr = std::shared_ptr<sinks::sink>;
}
""")
    src = tmp_path / "user.cpp"
    src.write_text(r"""
#include <cstdio>
#include <vector>
#include "tdcprocessor.h"
#ifndef SARLABTDC_AMD_HAVE_SPDLOG
#error "the spdlog overload is not declared"
#endif
struct capture : spdlog::sinks::sink {
    std::string last; spdlog::level::level_enum lvl = spdlog::level::off;
    void log(const spdlog::details::log_msg &m) override { last = m.logger_name + ":" + m.payload; lvl = m.level; }
    void flush() override {}
};
void (TdcProcessor::*overload)(const std::vector<spdlog::sink_ptr> &) = &TdcProcessor::setLoggerSinks;
int main() {
    auto c = std::make_shared<capture>();
    tdclog::spdlog_sink fwd(c);
    fwd.log(tdclog::level::warn, "TDCPROC", "Processed 8 chunks, skipped 0 chunks");
    std::printf("%s %d\n", c->last.c_str(), int(c->lvl));
    return (c->last == "TDCPROC:Processed 8 chunks, skipped 0 chunks" && c->lvl == spdlog::level::warn && overload) ? 0 : 1;
}
""")
    import subprocess

    pkg = os.path.join(ROOT, "sarlab-tdc_amd", "sarlabtdc")
    exe = tmp_path / "user"
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
           "-I", str(tmp_path / "inc"), "-I", os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           str(src), "-L", pkg, "-lsarlabtdc_amd", "-Wl,-rpath," + pkg, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True)
    assert run.returncode == 0, run.stdout + run.stderr
