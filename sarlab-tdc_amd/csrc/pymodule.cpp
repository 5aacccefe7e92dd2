// pymodule.cpp -- the _sarlabtdc Python module.
//
// Same surface as the reference module (py-sarlabtdc/src/interface.cpp:9-23,
// py-sarlabtdc/src/tdcprocessor_wrapper.cpp): class TdcProcessor(gpuNum) with
// start(dopplerWinCenter, dopplerBandwidth, dopCentroidWin, applyRangeWin,
// dopplerWinTaper=True), positional setRawData(...), setFocusGrid(grid) and
// getFocusedImage(). Arrays are py::array_t<T, c_style> without forcecast,
// as in tdcprocessor_wrapper.h:27-35, and the ndim / shape checks raise the
// same messages. Differences: the GIL is released while the GPU works, and
// getFocusedImage() before start() raises instead of reading a null buffer.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "tdcprocessor.h"

namespace py = pybind11;

namespace {

using c64_array = py::array_t<std::complex<float>, py::array::c_style>;
using f32_array = py::array_t<float, py::array::c_style>;
using f64_array = py::array_t<double, py::array::c_style>;

int parse_form(const py::object &o)
{
    if (py::isinstance<py::str>(o)) {
        const auto s = o.cast<std::string>();
        if (s == "plain") return TDC_FORM_PLAIN;
        if (s == "precision") return TDC_FORM_PRECISION;
        throw std::invalid_argument("form must be 'plain' or 'precision'");
    }
    return o.cast<int>();
}

int parse_storage(const py::object &o)
{
    if (py::isinstance<py::str>(o)) {
        const auto s = o.cast<std::string>();
        if (s == "c64" || s == "complex64") return TDC_RAW_C64;
        if (s == "f16" || s == "float16") return TDC_RAW_F16;
        if (s == "bf16" || s == "bfloat16") return TDC_RAW_BF16;
        throw std::invalid_argument("storage must be 'c64', 'f16' or 'bf16'");
    }
    return o.cast<int>();
}

int parse_window(const py::object &o)
{
    if (py::isinstance<py::str>(o)) {
        const auto s = o.cast<std::string>();
        if (s == "hamming") return TDC_WIN_HAMMING;
        if (s == "hann") return TDC_WIN_HANN;
        if (s == "taylor") return TDC_WIN_TAYLOR;
        throw std::invalid_argument("window must be 'hamming', 'hann' or 'taylor'");
    }
    return o.cast<int>();
}

class PyTdcProcessor
{
public:
    explicit PyTdcProcessor(int gpuNum) : proc_(std::make_unique<TdcProcessor>(gpuNum))
    {
        // the reference wrapper always logs to stdout (tdcprocessor_wrapper.cpp:165-170)
        proc_->setLoggerSinks({std::make_shared<tdclog::stdout_sink>()});
    }

    void start(float dopplerWinCenter, float dopplerBandwidth, bool dopCentroidWin,
               bool applyRangeWin, bool dopplerWinTaper)
    {
        py::gil_scoped_release nogil;
        proc_->start(dopplerWinCenter, dopplerBandwidth, dopCentroidWin, applyRangeWin,
                     dopplerWinTaper);
        started_ = true;
    }

    void setRawData(c64_array rawData, f32_array priTimes, f32_array sampleTimes,
                    f32_array position, f32_array velocity, f32_array attitude, float modRate,
                    float startFreq, f32_array bodyBoresight)
    {
        // tdcprocessor_wrapper.cpp:36-113
        auto di = rawData.request(), pi = priTimes.request(), si = sampleTimes.request(),
             po = position.request(), ve = velocity.request(), at = attitude.request(),
             bo = bodyBoresight.request();
        if (di.ndim != 2) throw std::runtime_error("rawData must be 2D");
        if (pi.ndim != 1) throw std::runtime_error("priTimes must be 1D");
        if (si.ndim != 1) throw std::runtime_error("sampleTimes must be 1D");
        if (po.ndim != 3) throw std::runtime_error("position must be 3D");
        if (ve.ndim != 3) throw std::runtime_error("velocity must be 3D");
        if (at.ndim != 3) throw std::runtime_error("attitude must be 3D");
        if (bo.ndim != 1) throw std::runtime_error("bodyBoresight must be 1D");
        const auto nPri = static_cast<int>(di.shape[0]);
        const auto nSamples = static_cast<int>(di.shape[1]);
        if (pi.shape[0] != nPri) throw std::runtime_error("priTimes shape is incorrect");
        if (si.shape[0] != nSamples) throw std::runtime_error("sampleTimes shape is incorrect");
        if (po.shape[0] != nPri || po.shape[1] != nSamples || po.shape[2] != 3)
            throw std::runtime_error("position shape is incorrect");
        if (ve.shape[0] != nPri || ve.shape[1] != nSamples || ve.shape[2] != 3)
            throw std::runtime_error("velocity shape is incorrect");
        if (at.shape[0] != nPri || at.shape[1] != nSamples || at.shape[2] != 4)
            throw std::runtime_error("attitude shape is incorrect");
        if (bo.shape[0] != 3) throw std::runtime_error("bodyBoresight shape is incorrect");
        const auto *b = static_cast<const float *>(bo.ptr);
        float3 bore;
        bore.x = b[0];
        bore.y = b[1];
        bore.z = b[2];
        proc_->setRawData(static_cast<const std::complex<float> *>(di.ptr),
                          static_cast<const float *>(pi.ptr), static_cast<const float *>(si.ptr),
                          static_cast<const float *>(po.ptr), static_cast<const float *>(ve.ptr),
                          static_cast<const float *>(at.ptr), nPri, nSamples, modRate, startFreq,
                          bore);
        // keep the arrays alive while the processor borrows them (wrapper.h:45-53)
        keep_ = py::make_tuple(rawData, priTimes, sampleTimes, position, velocity, attitude,
                               bodyBoresight);
    }

    void setFocusGrid(f32_array focusGrid)
    {
        auto gi = focusGrid.request();
        if (gi.ndim != 3) throw std::runtime_error("focusGrid must be 3D");
        const auto rows = static_cast<int>(gi.shape[0]);
        const auto cols = static_cast<int>(gi.shape[1]);
        if (gi.shape[2] != 3) throw std::runtime_error("focusGrid shape is incorrect");
        proc_->setFocusGrid(static_cast<const float *>(gi.ptr), rows, cols);
        keep_grid_ = focusGrid;
    }

    c64_array getFocusedImage()
    {
        const float2 *img = proc_->imageBuffer();
        if (!img) throw std::runtime_error("no focused image: call start() first");
        const int rows = proc_->rows(), cols = proc_->cols();
        c64_array out({rows, cols});
        std::memcpy(out.mutable_data(), img,
                    static_cast<size_t>(rows) * cols * sizeof(std::complex<float>));
        return out;
    }

    // ---- additive API ----
    void setRawDataF64(c64_array rawData, f64_array priTimes, f64_array sampleTimes,
                       f64_array position, double modRate, double startFreq)
    {
        auto di = rawData.request(), pi = priTimes.request(), si = sampleTimes.request(),
             po = position.request();
        if (di.ndim != 2) throw std::runtime_error("rawData must be 2D");
        if (pi.ndim != 1) throw std::runtime_error("priTimes must be 1D");
        if (si.ndim != 1) throw std::runtime_error("sampleTimes must be 1D");
        if (po.ndim != 2) throw std::runtime_error("position must be 2D (one phase centre per chirp)");
        const auto nPri = static_cast<int>(di.shape[0]);
        const auto nSamples = static_cast<int>(di.shape[1]);
        if (pi.shape[0] != nPri) throw std::runtime_error("priTimes shape is incorrect");
        if (si.shape[0] != nSamples) throw std::runtime_error("sampleTimes shape is incorrect");
        if (po.shape[0] != nPri || po.shape[1] != 3) throw std::runtime_error("position shape is incorrect");
        proc_->setRawDataF64(static_cast<const std::complex<float> *>(di.ptr),
                             static_cast<const double *>(pi.ptr), static_cast<const double *>(si.ptr),
                             static_cast<const double *>(po.ptr), nPri, nSamples, modRate, startFreq);
        keep_ = py::make_tuple(rawData, priTimes, sampleTimes, position);
    }

    void setFocusGridF64(f64_array focusGrid)
    {
        auto gi = focusGrid.request();
        if (gi.ndim != 3) throw std::runtime_error("focusGrid must be 3D");
        if (gi.shape[2] != 3) throw std::runtime_error("focusGrid shape is incorrect");
        proc_->setFocusGridF64(static_cast<const double *>(gi.ptr), static_cast<int>(gi.shape[0]),
                               static_cast<int>(gi.shape[1]));
        keep_grid_ = focusGrid;
    }

    void setForm(const py::object &form) { proc_->setForm(parse_form(form)); }
    void setRawStorage(const py::object &s) { proc_->setRawStorage(parse_storage(s)); }
    void setRangeWindow(const py::object &kind, double p0, double p1)
    {
        int k = parse_window(kind);
        if (k == TDC_WIN_HAMMING && p0 == 0.0) p0 = 25.0 / 46.0;
        proc_->setRangeWindow(k, p0, p1);
    }
    void setLaunch(int ppl, int splits) { proc_->setLaunch(ppl, splits); }
    void setLogging(bool on)
    {
        if (on) proc_->setLoggerSinks({std::make_shared<tdclog::stdout_sink>()});
        else proc_->setLoggerSinks({});
    }
    void prepare(float c, float bw, bool centroid, bool rw, bool taper)
    {
        py::gil_scoped_release nogil;
        proc_->prepare(c, bw, centroid, rw, taper);
    }
    void run(int iterations)
    {
        py::gil_scoped_release nogil;
        proc_->run(iterations);
    }
    void fetchImage()
    {
        py::gil_scoped_release nogil;
        proc_->fetchImage();
        started_ = true;
    }
    py::dict stats() const
    {
        const tdc_stats st = proc_->stats();
        py::dict d;
        d["kernel_ms"] = st.kernel_ms;
        d["main_kernel_ms"] = st.main_kernel_ms;
        d["end_to_end_ms"] = st.end_to_end_ms;
        d["terms"] = st.terms;
        d["form"] = st.form == TDC_FORM_PLAIN ? "plain" : "precision";
        d["raw_storage"] = st.raw_storage == TDC_RAW_C64 ? "c64" : (st.raw_storage == TDC_RAW_F16 ? "f16" : "bf16");
        d["doppler_active"] = static_cast<bool>(st.doppler_active);
        d["chirp_splits"] = st.chirp_splits;
        d["pixels_per_lane"] = st.pixels_per_lane;
        d["workgroups"] = st.workgroups;
        d["skipped_chunks"] = st.skipped_chunks;
        d["processed_chunks"] = st.processed_chunks;
        return d;
    }

private:
    std::unique_ptr<TdcProcessor> proc_;
    py::object keep_ = py::none();
    py::object keep_grid_ = py::none();
    bool started_ = false;
};

} // namespace

PYBIND11_MODULE(_sarlabtdc, m)
{
    m.doc() = "Time-domain correlation SAR processor for high bandwidth mmWave radars (MI355X).";
    m.attr("FORM_PLAIN") = TDC_FORM_PLAIN;
    m.attr("FORM_PRECISION") = TDC_FORM_PRECISION;
    m.def("device_count", &tdc_device_count, "Number of visible HIP devices.");

    py::class_<PyTdcProcessor>(m, "TdcProcessor")
        .def(py::init<int>())
        .def("start", &PyTdcProcessor::start, py::arg("dopplerWinCenter"),
             py::arg("dopplerBandwidth"), py::arg("dopCentroidWin"), py::arg("applyRangeWin"),
             py::arg("dopplerWinTaper") = true)
        .def("setRawData", &PyTdcProcessor::setRawData)
        .def("setFocusGrid", &PyTdcProcessor::setFocusGrid)
        .def("getFocusedImage", &PyTdcProcessor::getFocusedImage)
        .def("setRawDataF64", &PyTdcProcessor::setRawDataF64)
        .def("setFocusGridF64", &PyTdcProcessor::setFocusGridF64)
        .def("setForm", &PyTdcProcessor::setForm, py::arg("form"))
        .def("setRawStorage", &PyTdcProcessor::setRawStorage, py::arg("storage"))
        .def("setRangeWindow", &PyTdcProcessor::setRangeWindow, py::arg("kind"),
             py::arg("p0") = 0.0, py::arg("p1") = 0.0)
        .def("setLaunch", &PyTdcProcessor::setLaunch, py::arg("pixelsPerLane"),
             py::arg("chirpSplits") = 0)
        .def("setLogging", &PyTdcProcessor::setLogging, py::arg("enabled"))
        .def("prepare", &PyTdcProcessor::prepare, py::arg("dopplerWinCenter"),
             py::arg("dopplerBandwidth"), py::arg("dopCentroidWin"), py::arg("applyRangeWin"),
             py::arg("dopplerWinTaper") = true)
        .def("run", &PyTdcProcessor::run, py::arg("iterations") = 1)
        .def("fetchImage", &PyTdcProcessor::fetchImage)
        .def("stats", &PyTdcProcessor::stats);
}
