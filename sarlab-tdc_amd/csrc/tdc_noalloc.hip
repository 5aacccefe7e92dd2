// tdc_noalloc.hip -- the allocation-free entry point tdc_start_noalloc.
//
// No counterpart in the reference. It forms the precision-form image of
// tdc_start (range-profile engine) with every device buffer carved out of
// one file-scope device array, `scratch`, of kNoallocScratchBytes per GPU,
// instead of the processor's growable buffers: for callers that cannot
// allocate device memory at the point of the call (serving loops, harnesses
// that forbid hipMalloc). The array lives as long as the process on every GPU
// the library runs on. Its device address differs per GPU and is looked up
// once per device with hipGetSymbolAddress; host code never passes the
// array's own name to a kernel or a copy (in host code that name is a host
// stand-in, not a device address).
//
// With the same chirp split count the image is bit-identical to tdc_start's:
// the same host tables, the same staging, correlate-sum and split reduction
// kernels. The split count is tdc_start's automatic one, halved until the
// buffers fit the array (a count forced with tdc_set_launch must fit as is).
#include <chrono>
#include <mutex>
#include <string>

#include <hip/hip_runtime.h>

#include "sarlabtdc_amd.h"
#include "tdc_internal.h"

namespace {

constexpr size_t kNoallocScratchBytes = size_t(TDC_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;

} // namespace

// Device buffers of tdc_start_noalloc: raw samples, range window, kernel
// correction, twiddles, chirp records, grid, range-profile tables, partial
// images and the image, at 256-byte aligned offsets.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_mu;                       // guards g_addr
void *g_addr[kMaxDevices] = {};        // the array's device address per device
std::mutex g_use[kMaxDevices];         // serialises the calls on one device

// The array's address on `device`, which must be current.
int scratch_address(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return tdc::set_last_error(TDC_ERR_DEVICE,
                                       std::string("hipGetSymbolAddress(scratch) failed: ") +
                                           hipGetErrorString(e));
        g_addr[device] = p;
    }
    *out = static_cast<unsigned char *>(g_addr[device]);
    return TDC_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Offsets of the buffers in the array for `nsplit` partial images.
struct Layout {
    size_t raw, wr, corr, tw, pcpos, grid, table, part, img, end;
};

Layout layout(const tdc::NoallocInputs &in, int nsplit)
{
    const size_t nterm = static_cast<size_t>(in.npri) * in.ns;
    const size_t npx = static_cast<size_t>(in.npx);
    Layout l{};
    size_t o = 0;
    l.raw = o;
    o += align256(sizeof(float2) * nterm);
    l.wr = o;
    o += align256(sizeof(float) * in.ns);
    l.corr = o;
    o += align256(sizeof(float) * in.tabs->corr.size());
    l.tw = o;
    o += align256(sizeof(float2) * in.tabs->twiddle.size());
    l.pcpos = o;
    o += align256(sizeof(tdc::ChirpPos) * in.npri);
    l.grid = o;
    o += align256(3 * sizeof(double) * npx);
    l.table = o;
    o += align256(sizeof(float2) * static_cast<size_t>(in.npri) * in.pp.Mp);
    l.part = o;
    o += align256(sizeof(float2) * npx * nsplit); // FP32 partial images, one per split
    l.img = o;
    o += align256(sizeof(float2) * npx);
    l.end = o;
    return l;
}

#define NA_HIP(expr)                                                                         \
    do {                                                                                     \
        const hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                                \
            return tdc::set_last_error(TDC_ERR_DEVICE,                                      \
                                       std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

} // namespace

extern "C" int tdc_start_noalloc(tdc_processor *h, float doppler_win_center,
                                 float doppler_bandwidth, int dop_centroid_win,
                                 int apply_range_win, int doppler_win_taper, float *image_iq,
                                 size_t image_floats)
{
    const auto t0 = std::chrono::steady_clock::now();
    tdc::Range range("tdc_start_noalloc");
    tdc::NoallocInputs in;
    int rc = tdc::noalloc_inputs(h, doppler_win_center, doppler_bandwidth, dop_centroid_win != 0,
                                 apply_range_win != 0, doppler_win_taper != 0, in);
    if (rc != TDC_OK) return rc;
    if (in.npx > 0 && !image_iq)
        return tdc::set_last_error(TDC_ERR_INVALID_ARGUMENT, "image_iq is NULL");
    if (image_floats < 2 * static_cast<size_t>(in.npx))
        return tdc::set_last_error(TDC_ERR_INVALID_ARGUMENT,
                                   "image_iq holds " + std::to_string(image_floats) +
                                       " floats, the image needs " +
                                       std::to_string(2 * static_cast<size_t>(in.npx)));
    // the split count: tdc_start's, halved until the buffers fit the array
    int nsplit = in.nsplit;
    Layout l = layout(in, nsplit);
    while (!in.forced && l.end > kNoallocScratchBytes && nsplit > 1) {
        nsplit /= 2;
        l = layout(in, nsplit);
    }
    if (l.end > kNoallocScratchBytes)
        return tdc::set_last_error(
            TDC_ERR_INVALID_ARGUMENT,
            "the image needs " + std::to_string(l.end) + " bytes of scratch with " +
                std::to_string(nsplit) + " chirp split(s); the allocation-free entry point has " +
                std::to_string(kNoallocScratchBytes) + " (TDC_NOALLOC_SCRATCH_BYTES)");

    if (in.device < 0 || in.device >= kMaxDevices)
        return tdc::set_last_error(TDC_ERR_DEVICE, "device index beyond the supported range");
    std::lock_guard<std::mutex> use(g_use[in.device]);
    unsigned char *base = nullptr;
    if ((rc = scratch_address(in.device, &base)) != TDC_OK) return rc;
    const hipStream_t s = in.stream;
    const size_t nterm = static_cast<size_t>(in.npri) * in.ns;
    const size_t npx = static_cast<size_t>(in.npx);
    auto *raw = reinterpret_cast<float *>(base + l.raw);
    auto *wr = reinterpret_cast<float *>(base + l.wr);
    auto *corr = reinterpret_cast<float *>(base + l.corr);
    auto *tw = reinterpret_cast<float2 *>(base + l.tw);
    auto *pcpos = reinterpret_cast<tdc::ChirpPos *>(base + l.pcpos);
    auto *grid = reinterpret_cast<double *>(base + l.grid);
    auto *table = reinterpret_cast<float2 *>(base + l.table);
    auto *part = reinterpret_cast<float2 *>(base + l.part);
    auto *img = reinterpret_cast<float2 *>(base + l.img);

    // inputs from host memory (copies through the runtime's own staging)
    NA_HIP(hipMemcpyAsync(raw, in.raw, sizeof(float2) * nterm, hipMemcpyHostToDevice, s));
    NA_HIP(hipMemcpyAsync(wr, in.wr.data(), sizeof(float) * in.ns, hipMemcpyHostToDevice, s));
    NA_HIP(hipMemcpyAsync(corr, in.tabs->corr.data(), sizeof(float) * in.tabs->corr.size(),
                          hipMemcpyHostToDevice, s));
    NA_HIP(hipMemcpyAsync(tw, in.tabs->twiddle.data(), sizeof(float2) * in.tabs->twiddle.size(),
                          hipMemcpyHostToDevice, s));
    NA_HIP(hipMemcpyAsync(pcpos, in.pc.data(), sizeof(tdc::ChirpPos) * in.pc.size(),
                          hipMemcpyHostToDevice, s));
    if (npx)
        NA_HIP(hipMemcpyAsync(grid, in.grid, 3 * sizeof(double) * npx, hipMemcpyHostToDevice, s));

    // the kernels of tdc_start's precision pass: per-chirp tables (window and
    // storage rounding folded in), correlate-sum, fixed-order split reduction
    tdc::Geom g{};
    tdc::make_launch_geom(g, in.npx, in.npri, in.ns, 1, nsplit);
    int wgs = 0;
    NA_HIP(tdc::launch_stage_profile_folded(in.storage, raw, wr, corr, tw, pcpos, in.npri, in.ns,
                                            in.pp.M, in.pp.Mp, in.pp.taps, in.pp.reverse, table,
                                            s));
    if (npx) {
        NA_HIP(tdc::launch_profile(table, pcpos, grid, in.pp, g, part, s, &wgs));
        NA_HIP(tdc::launch_finalize_p32(part, nsplit, in.npx, img, s));
        NA_HIP(hipMemcpyAsync(image_iq, img, sizeof(float2) * npx, hipMemcpyDeviceToHost, s));
    }
    NA_HIP(hipStreamSynchronize(s));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    tdc::noalloc_finish(h, in, nsplit, wgs, ms);
    return TDC_OK;
}
