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
#include <map>
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

int parse_engine(const py::object &o)
{
    if (py::isinstance<py::str>(o)) {
        const auto s = o.cast<std::string>();
        if (s == "auto") return TDC_ENGINE_AUTO;
        if (s == "vector") return TDC_ENGINE_VECTOR;
        if (s == "matrix") return TDC_ENGINE_MATRIX;
        if (s == "profile") return TDC_ENGINE_PROFILE;
        throw std::invalid_argument("engine must be 'auto', 'vector' or 'matrix'");
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
    void setEngine(const py::object &e) { proc_->setEngine(parse_engine(e)); }
    void setProfileKernel(int taps, int over) { proc_->setProfileKernel(taps, over); }
    void setProfileExactRange(bool exact) { proc_->setProfileExactRange(exact); }
    void setRangeWindow(const py::object &kind, double p0, double p1)
    {
        int k = parse_window(kind);
        if (k == TDC_WIN_HAMMING && p0 == 0.0) p0 = 25.0 / 46.0;
        proc_->setRangeWindow(k, p0, p1);
    }
    void setLaunch(int ppl, int splits) { proc_->setLaunch(ppl, splits); }
    void setDeviceBudget(unsigned long long bytes) { proc_->setDeviceBudget(bytes); }
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
        d["range_per_chirp"] = (st.per_chirp & TDC_PER_CHIRP_RANGE) != 0;
        d["window_per_chirp"] = (st.per_chirp & TDC_PER_CHIRP_WINDOW) != 0;
        d["range_series"] = (st.per_chirp & TDC_PER_TERM_SERIES) != 0;
        d["engine"] = st.engine == TDC_ENGINE_MATRIX    ? "matrix"
                      : st.engine == TDC_ENGINE_VECTOR  ? "vector"
                      : st.engine == TDC_ENGINE_PROFILE ? "profile"
                                                        : "none";
        d["taps"] = st.profile_taps;
        d["oversampling"] = st.profile_oversampling;
        d["incremental"] = st.profile_incremental != 0;
        return d;
    }

private:
    std::unique_ptr<TdcProcessor> proc_;
    py::object keep_ = py::none();
    py::object keep_grid_ = py::none();
    bool started_ = false;
};

void check_status(int rc)
{
    if (rc == TDC_OK) return;
    if (rc == TDC_ERR_INVALID_ARGUMENT) throw std::invalid_argument(tdc_last_error());
    throw std::runtime_error(tdc_last_error());
}

// Serving form: continuously batched jobs on one GPU.
class PyServer
{
public:
    PyServer(int gpuNum, double hbmBudgetBytes, double stepTerms)
    {
        check_status(tdc_server_create(gpuNum, hbmBudgetBytes, stepTerms, &s_));
    }
    ~PyServer() { close(); }

    int64_t submit(c64_array rawData, f64_array sampleTimes, f64_array position, f64_array focusGrid,
                   double modRate, double startFreq, const py::object &form,
                   const py::object &storage)
    {
        if (!s_) throw std::runtime_error("the server is closed");
        auto di = rawData.request(), si = sampleTimes.request(), po = position.request(),
             gi = focusGrid.request();
        if (di.ndim != 2) throw std::runtime_error("rawData must be 2D");
        if (si.ndim != 1 || si.shape[0] != di.shape[1])
            throw std::runtime_error("sampleTimes shape is incorrect");
        if (po.ndim != 2 || po.shape[0] != di.shape[0] || po.shape[1] != 3)
            throw std::runtime_error("position shape is incorrect");
        if (gi.ndim != 3 || gi.shape[2] != 3) throw std::runtime_error("focusGrid shape is incorrect");
        tdc_job_desc d{};
        d.raw_iq = static_cast<const float *>(di.ptr);
        d.sample_times = static_cast<const double *>(si.ptr);
        d.chirp_position = static_cast<const double *>(po.ptr);
        d.grid = static_cast<const double *>(gi.ptr);
        d.n_pri = static_cast<int32_t>(di.shape[0]);
        d.n_samples = static_cast<int32_t>(di.shape[1]);
        d.n_rows = static_cast<int32_t>(gi.shape[0]);
        d.n_cols = static_cast<int32_t>(gi.shape[1]);
        d.mod_rate = modRate;
        d.start_freq = startFreq;
        d.form = parse_form(form);
        d.raw_storage = parse_storage(storage);
        int64_t id = 0;
        check_status(tdc_server_submit(s_, &d, &id));
        keep_[id] = Keep{py::make_tuple(rawData, sampleTimes, position, focusGrid), d.n_rows, d.n_cols};
        return id;
    }

    bool poll(int64_t id)
    {
        int done = 0;
        check_status(tdc_server_poll(s_, id, &done));
        return done != 0;
    }

    py::tuple wait(int64_t id)
    {
        auto it = keep_.find(id);
        if (it == keep_.end()) throw std::invalid_argument("unknown job id");
        c64_array img({it->second.rows, it->second.cols});
        tdc_job_info info{};
        int rc;
        {
            float *dst = reinterpret_cast<float *>(img.mutable_data());
            py::gil_scoped_release nogil;
            rc = tdc_server_wait(s_, id, dst, &info);
        }
        keep_.erase(it);
        check_status(rc);
        py::dict d;
        d["id"] = info.id;
        d["t_submit"] = info.t_submit_s;
        d["t_admit"] = info.t_admit_s;
        d["t_first_step"] = info.t_first_step_s;
        d["t_done"] = info.t_done_s;
        d["latency_ms"] = (info.t_done_s - info.t_submit_s) * 1e3;
        d["terms"] = info.terms;
        d["device_bytes"] = info.device_bytes;
        d["tiles"] = info.tiles;
        d["steps"] = info.steps;
        return py::make_tuple(img, d);
    }

    py::dict stats()
    {
        tdc_server_stats st{};
        check_status(tdc_server_get_stats(s_, &st));
        py::dict d;
        d["submitted"] = st.submitted;
        d["completed"] = st.completed;
        d["failed"] = st.failed;
        d["steps"] = st.steps;
        d["device_busy_ms"] = st.device_busy_ms;
        d["terms_done"] = st.terms_done;
        d["bytes_in_use"] = st.bytes_in_use;
        d["bytes_peak"] = st.bytes_peak;
        d["queued"] = st.queued;
        d["active"] = st.active;
        d["graphs"] = st.graphs;
        return d;
    }

    double clock()
    {
        double t = 0.0;
        check_status(tdc_server_clock(s_, &t));
        return t;
    }

    void close()
    {
        if (!s_) return;
        tdc_server *s = s_;
        s_ = nullptr;
        py::gil_scoped_release nogil;
        tdc_server_destroy(s);
    }

private:
    struct Keep {
        py::object arrays;
        int rows, cols;
    };
    tdc_server *s_ = nullptr;
    std::map<int64_t, Keep> keep_;
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
        .def("setEngine", &PyTdcProcessor::setEngine, py::arg("engine"))
        .def("setProfileKernel", &PyTdcProcessor::setProfileKernel, py::arg("taps") = 0,
             py::arg("oversampling") = 0)
        .def("setProfileExactRange", &PyTdcProcessor::setProfileExactRange, py::arg("exact") = true)
        .def("setRangeWindow", &PyTdcProcessor::setRangeWindow, py::arg("kind"),
             py::arg("p0") = 0.0, py::arg("p1") = 0.0)
        .def("setLaunch", &PyTdcProcessor::setLaunch, py::arg("pixelsPerLane"),
             py::arg("chirpSplits") = 0)
        .def("setDeviceBudget", &PyTdcProcessor::setDeviceBudget, py::arg("bytes"))
        .def("setLogging", &PyTdcProcessor::setLogging, py::arg("enabled"))
        .def("prepare", &PyTdcProcessor::prepare, py::arg("dopplerWinCenter"),
             py::arg("dopplerBandwidth"), py::arg("dopCentroidWin"), py::arg("applyRangeWin"),
             py::arg("dopplerWinTaper") = true)
        .def("run", &PyTdcProcessor::run, py::arg("iterations") = 1)
        .def("fetchImage", &PyTdcProcessor::fetchImage)
        .def("stats", &PyTdcProcessor::stats);

    py::class_<PyServer>(m, "Server")
        .def(py::init<int, double, double>(), py::arg("gpuNum") = 0,
             py::arg("hbmBudgetBytes") = 200e9, py::arg("stepTerms") = 0.0)
        .def("submit", &PyServer::submit, py::arg("rawData"), py::arg("sampleTimes"),
             py::arg("position"), py::arg("focusGrid"), py::arg("modRate"), py::arg("startFreq"),
             py::arg("form") = "precision", py::arg("storage") = "c64")
        .def("poll", &PyServer::poll, py::arg("jobId"))
        .def("wait", &PyServer::wait, py::arg("jobId"))
        .def("stats", &PyServer::stats)
        .def("clock", &PyServer::clock)
        .def("close", &PyServer::close);
}
