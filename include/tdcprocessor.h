/*
 * tdcprocessor.h -- C++ interface of the MI355X TDC SAR processor.
 *
 * Same public surface as the reference class TdcProcessor
 * (tdc-lib/include/tdcprocessor.h:42-135): constructor(gpuNum), start(),
 * setRawData(), setFocusGrid(), setLoggerSinks() and imageBuffer(), with the
 * same argument order, types and defaults, and the same exceptions
 * (std::invalid_argument for nPri/nSamples <= 0, std::runtime_error for
 * device and processing failures). It is a thin layer over the C ABI in
 * sarlabtdc_amd.h.
 *
 * Usage (as the reference):
 *   1) create an instance, 2) setRawData() and setFocusGrid(), 3) start(),
 *   4) imageBuffer() for the focused image (row-major [nRows][nCols]).
 *
 * Coordinates: the focus grid and the antenna positions share one local
 * frame; attitude quaternions (scalar first) rotate the body frame into it.
 *
 * Logging: spdlog is not a dependency of the library. setLoggerSinks() takes
 * sinks of the minimal interface tdclog::sink below; when the including
 * translation unit can see <spdlog/spdlog.h>, the reference's overload taking
 * std::vector<spdlog::sink_ptr> (tdcprocessor.h:128) is also declared, inline,
 * forwarding each line to the spdlog sinks (an empty braced list {} is then
 * ambiguous between the two overloads: pass a typed vector).
 */
#ifndef SARLABTDC_AMD_TDCPROCESSOR_H
#define SARLABTDC_AMD_TDCPROCESSOR_H

#include <complex>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_vector_types.h>

#include "sarlabtdc_amd.h"

namespace tdclog {

enum class level { debug, info, warn, error };

/* A log destination. */
class sink
{
public:
    virtual ~sink() = default;
    virtual void log(level lvl, const std::string &logger, const std::string &msg) = 0;
};

using sink_ptr = std::shared_ptr<sink>;

/* Writes "[time] [logger] [level] message" lines to stdout. */
class stdout_sink : public sink
{
public:
    void log(level lvl, const std::string &logger, const std::string &msg) override;
};

} // namespace tdclog

#if defined(__has_include)
#if __has_include(<spdlog/spdlog.h>)
#include <spdlog/spdlog.h>
#define SARLABTDC_AMD_HAVE_SPDLOG 1
namespace tdclog {
/* Forwards the processor's lines to one spdlog sink. */
class spdlog_sink : public sink
{
public:
    explicit spdlog_sink(spdlog::sink_ptr target) : target_(std::move(target)) {}
    void log(level lvl, const std::string &logger, const std::string &msg) override
    {
        const spdlog::level::level_enum l = lvl == level::debug  ? spdlog::level::debug
                                            : lvl == level::info ? spdlog::level::info
                                            : lvl == level::warn ? spdlog::level::warn
                                                                 : spdlog::level::err;
        target_->log(spdlog::details::log_msg(logger, l, msg));
    }

private:
    spdlog::sink_ptr target_;
};
} // namespace tdclog
#endif
#endif

class TdcProcessor
{
public:
    explicit TdcProcessor(int gpuNum);
    ~TdcProcessor();
    TdcProcessor(const TdcProcessor &) = delete;
    TdcProcessor &operator=(const TdcProcessor &) = delete;

    /* tdcprocessor.h:67-68 */
    void start(float dopplerWinCenter, float dopplerBandwidth, bool dopCentroidWin,
               bool applyRangeWin, bool dopplerWinTaper = true);

    /* tdcprocessor.h:104-107 */
    void setRawData(std::complex<float> const *rawData, float const *priTimes,
                    float const *sampleTimes, float const *position, float const *velocity,
                    float const *attitude, int nPri, int nSamples, float modRate,
                    float startFreq, float3 bodyBoresight);

    /* tdcprocessor.h:121 */
    void setFocusGrid(float const *focusGrid, int nRows, int nCols);

    /* tdcprocessor.h:128 (sink type: see the note at the top of this file) */
    void setLoggerSinks(const std::vector<tdclog::sink_ptr> &sinks);
#ifdef SARLABTDC_AMD_HAVE_SPDLOG
    /* tdcprocessor.h:128 with the reference's spdlog sinks */
    void setLoggerSinks(const std::vector<spdlog::sink_ptr> &sinks)
    {
        std::vector<tdclog::sink_ptr> wrapped;
        for (const spdlog::sink_ptr &s : sinks)
            wrapped.push_back(std::make_shared<tdclog::spdlog_sink>(s));
        setLoggerSinks(wrapped);
    }
#endif

    /* tdcprocessor.h:135: nullptr before the first start() */
    float2 const *imageBuffer() const;

    /* ---- additive API: no counterpart in the reference ---- */

    /* FP64 inputs for the precision form (chirpPosition is [nPri][3]). */
    void setRawDataF64(std::complex<float> const *rawData, double const *priTimes,
                       double const *sampleTimes, double const *chirpPosition, int nPri,
                       int nSamples, double modRate, double startFreq);
    void setFocusGridF64(double const *focusGrid, int nRows, int nCols);
    void setForm(int form);             /* TDC_FORM_* */
    void setRawStorage(int storage);    /* TDC_RAW_* */
    void setEngine(int engine);         /* TDC_ENGINE_* (precision form) */
    void setProfileKernel(int taps, int oversampling); /* range-profile kernel (0, 0: auto) */
    void setProfileExactRange(bool exact);              /* FP64 range for every chirp */
    void setRangeWindow(int kind, double p0 = 0.0, double p1 = 0.0); /* TDC_WIN_* */
    void setLaunch(int pixelsPerLane, int chirpSplits);
    void setDeviceBudget(unsigned long long bytes); /* 0: no limit */
    void prepare(float dopplerWinCenter, float dopplerBandwidth, bool dopCentroidWin,
                 bool applyRangeWin, bool dopplerWinTaper = true);
    void run(int iterations = 1);
    void fetchImage();
    tdc_stats stats() const;
    int rows() const;
    int cols() const;

private:
    void check(int status) const;
    void info(const std::string &msg) const;

    tdc_processor *h_ = nullptr;
    std::vector<tdclog::sink_ptr> sinks_;
};

#endif /* SARLABTDC_AMD_TDCPROCESSOR_H */
