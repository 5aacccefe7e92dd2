// tdcprocessor.cpp -- the C++ TdcProcessor class over the C ABI.
// Mirrors the reference class's behaviour (tdc-lib/src/tdcprocessor.cpp):
// status codes become std::invalid_argument / std::runtime_error, and the
// log lines follow the reference's progress messages.
#include "tdcprocessor.h"

#include <chrono>
#include <cstdio>
#include <ctime>
#include <stdexcept>

namespace tdclog {

void stdout_sink::log(level lvl, const std::string &logger, const std::string &msg)
{
    static const char *names[] = {"debug", "info", "warning", "error"};
    const auto now = std::chrono::system_clock::now();
    const std::time_t t = std::chrono::system_clock::to_time_t(now);
    const int ms = static_cast<int>(
        std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()).count() %
        1000);
    std::tm tmv{};
    localtime_r(&t, &tmv);
    char stamp[32];
    std::strftime(stamp, sizeof stamp, "%Y-%m-%d %H:%M:%S", &tmv);
    std::printf("[%s.%03d] [%s] [%s] %s\n", stamp, ms, logger.c_str(),
                names[static_cast<int>(lvl)], msg.c_str());
    std::fflush(stdout);
}

} // namespace tdclog

TdcProcessor::TdcProcessor(int gpuNum)
{
    tdc_processor *h = nullptr;
    const int rc = tdc_create(gpuNum, &h);
    if (rc != TDC_OK) throw std::runtime_error(tdc_last_error());
    h_ = h;
}

TdcProcessor::~TdcProcessor() { tdc_destroy(h_); }

void TdcProcessor::check(int status) const
{
    if (status == TDC_OK) return;
    if (status == TDC_ERR_INVALID_ARGUMENT) throw std::invalid_argument(tdc_last_error());
    throw std::runtime_error(tdc_last_error());
}

void TdcProcessor::info(const std::string &msg) const
{
    for (const auto &s : sinks_)
        if (s) s->log(tdclog::level::info, "TDCPROC", msg);
}

void TdcProcessor::start(float dopplerWinCenter, float dopplerBandwidth, bool dopCentroidWin,
                         bool applyRangeWin, bool dopplerWinTaper)
{
    info("Starting the TDC processor");
    check(tdc_start(h_, dopplerWinCenter, dopplerBandwidth, dopCentroidWin ? 1 : 0,
                    applyRangeWin ? 1 : 0, dopplerWinTaper ? 1 : 0));
    const tdc_stats st = stats();
    info("Processed " + std::to_string(st.processed_chunks) + " chunks, skipped " +
         std::to_string(st.skipped_chunks) + " chunks");
    info("Completed SAR processing");
    info("Elapsed time: " + std::to_string(st.end_to_end_ms / 1000.0) + "s");
}

void TdcProcessor::setRawData(std::complex<float> const *rawData, float const *priTimes,
                              float const *sampleTimes, float const *position,
                              float const *velocity, float const *attitude, int nPri,
                              int nSamples, float modRate, float startFreq, float3 bodyBoresight)
{
    info("Setting raw data with " + std::to_string(nPri) + " PRIs and " +
         std::to_string(nSamples) + " samples/PRI");
    const float bore[3] = {bodyBoresight.x, bodyBoresight.y, bodyBoresight.z};
    check(tdc_set_raw_data(h_, reinterpret_cast<const float *>(rawData), priTimes, sampleTimes,
                           position, velocity, attitude, nPri, nSamples, modRate, startFreq,
                           bore));
}

void TdcProcessor::setFocusGrid(float const *focusGrid, int nRows, int nCols)
{
    info("Configure the focus grid with shape " + std::to_string(nRows) + " x " +
         std::to_string(nCols));
    check(tdc_set_focus_grid(h_, focusGrid, nRows, nCols));
}

void TdcProcessor::setLoggerSinks(const std::vector<tdclog::sink_ptr> &sinks)
{
    sinks_ = sinks;
    info("Completed logging setup");
}

float2 const *TdcProcessor::imageBuffer() const
{
    const float *p = nullptr;
    check(tdc_image_buffer(h_, &p));
    return reinterpret_cast<float2 const *>(p);
}

void TdcProcessor::setRawDataF64(std::complex<float> const *rawData, double const *priTimes,
                                 double const *sampleTimes, double const *chirpPosition, int nPri,
                                 int nSamples, double modRate, double startFreq)
{
    check(tdc_set_raw_data_f64(h_, reinterpret_cast<const float *>(rawData), priTimes, sampleTimes,
                               chirpPosition, nPri, nSamples, modRate, startFreq));
}

void TdcProcessor::setFocusGridF64(double const *focusGrid, int nRows, int nCols)
{
    check(tdc_set_focus_grid_f64(h_, focusGrid, nRows, nCols));
}

void TdcProcessor::setForm(int form) { check(tdc_set_form(h_, form)); }
void TdcProcessor::setRawStorage(int storage) { check(tdc_set_raw_storage(h_, storage)); }
void TdcProcessor::setEngine(int engine) { check(tdc_set_engine(h_, engine)); }
void TdcProcessor::setProfileKernel(int taps, int oversampling)
{
    check(tdc_set_profile_kernel(h_, taps, oversampling));
}
void TdcProcessor::setProfileExactRange(bool exact)
{
    check(tdc_set_profile_exact_range(h_, exact ? 1 : 0));
}
void TdcProcessor::setRangeWindow(int kind, double p0, double p1)
{
    check(tdc_set_range_window(h_, kind, p0, p1));
}
void TdcProcessor::setLaunch(int pixelsPerLane, int chirpSplits)
{
    check(tdc_set_launch(h_, pixelsPerLane, chirpSplits));
}
void TdcProcessor::setDeviceBudget(unsigned long long bytes)
{
    check(tdc_set_device_budget(h_, bytes));
}
void TdcProcessor::prepare(float dopplerWinCenter, float dopplerBandwidth, bool dopCentroidWin,
                           bool applyRangeWin, bool dopplerWinTaper)
{
    check(tdc_prepare(h_, dopplerWinCenter, dopplerBandwidth, dopCentroidWin ? 1 : 0,
                      applyRangeWin ? 1 : 0, dopplerWinTaper ? 1 : 0));
}
void TdcProcessor::run(int iterations) { check(tdc_run(h_, iterations)); }
void TdcProcessor::fetchImage() { check(tdc_fetch_image(h_)); }

tdc_stats TdcProcessor::stats() const
{
    tdc_stats st{};
    check(tdc_get_stats(h_, &st));
    return st;
}

int TdcProcessor::rows() const
{
    int r = 0, c = 0;
    check(tdc_image_shape(h_, &r, &c));
    return r;
}

int TdcProcessor::cols() const
{
    int r = 0, c = 0;
    check(tdc_image_shape(h_, &r, &c));
    return c;
}
