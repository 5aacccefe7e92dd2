// tdc_capi.cpp -- host side of the C ABI declared in include/sarlabtdc_amd.h.
//
// Replaces the reference's TdcProcessor driver (tdc-lib/src/tdcprocessor.cpp:
// 59-167): instead of staging 128-chirp chunks through pinned buffers and
// launching five kernels per (pixel, chunk) on the legacy default stream,
// the whole job is uploaded once, staged by one kernel into the layout the
// correlate-sum kernel reads, formed by one correlate-sum launch (plus a
// fixed-order split reduction), and copied back. Device buffers are kept and
// reused across calls (the reference reallocates every buffer on every
// start(), tdcprocessor.cpp:65-66).
#include <algorithm>
#include <array>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include <hip/hip_runtime.h>

#include "sarlabtdc_amd.h"
#include "tdc_internal.h"

namespace {

constexpr double kSpeedOfLight = 299792458.0; // tdcprocessor.cpp:37
constexpr double kPi = 3.14159265358979323846;

thread_local std::string t_err;

int fail(int code, std::string msg)
{
    t_err = std::move(msg);
    return code;
}

int ok()
{
    t_err.clear();
    return TDC_OK;
}

#define TDC_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(TDC_ERR_DEVICE, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

// Page-locked host buffer, grown on demand: the image comes back from the
// device with one direct DMA (a copy into pageable memory is staged through
// the runtime's own pinned chunks, ~3x slower for a 2 MB image).
struct PinnedBuf {
    float *p = nullptr;
    size_t n = 0; // floats
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf()
    {
        if (p) (void) hipHostFree(p);
    }
    hipError_t ensure(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        if (p) (void) hipHostFree(p);
        p = nullptr;
        n = 0;
        void *q = nullptr;
        hipError_t e = hipHostMalloc(&q, sizeof(float) * std::max<size_t>(count, 64), hipHostMallocDefault);
        if (e != hipSuccess) return e;
        p = static_cast<float *>(q);
        n = std::max<size_t>(count, 64);
        return hipSuccess;
    }
};

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void) hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t ensure(size_t bytes)
    {
        if (bytes == 0) bytes = 256;
        if (bytes <= n) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Periodic (DFT-even) range windows, evaluated in FP64 and rounded once.
void range_window(int kind, double p0, double p1, int ns, bool apply, std::vector<float> &out)
{
    out.assign(ns, 1.0F);
    if (!apply) return;
    if (kind == TDC_WIN_HAMMING || kind == TDC_WIN_HANN) {
        // tdckernels.cu:43-46: A - (1-A) cos(2 pi m / Ns)
        const double a = (kind == TDC_WIN_HANN) ? 0.5 : p0;
        for (int m = 0; m < ns; ++m)
            out[m] = static_cast<float>(a - (1.0 - a) * std::cos(2.0 * kPi * m / ns));
        return;
    }
    // Taylor window (T. T. Taylor 1955), normalised to 1 at the centre,
    // periodic: the (ns+1)-point symmetric window without its last point.
    const int nbar = std::max(1, std::min(64, static_cast<int>(std::lround(p0))));
    const double sll = p1;
    const int nf = ns + 1;
    const double b = std::pow(10.0, sll / 20.0);
    const double a = std::acosh(b) / kPi;
    const double s2 = static_cast<double>(nbar) * nbar / (a * a + (nbar - 0.5) * (nbar - 0.5));
    std::vector<double> fm(std::max(0, nbar - 1));
    for (int m = 1; m < nbar; ++m) {
        double num = (m % 2 == 1) ? 1.0 : -1.0;
        double den = 2.0;
        for (int i = 1; i < nbar; ++i) {
            num *= 1.0 - static_cast<double>(m) * m / s2 / (a * a + (i - 0.5) * (i - 0.5));
            if (i != m) den *= 1.0 - static_cast<double>(m) * m / (static_cast<double>(i) * i);
        }
        fm[m - 1] = num / den;
    }
    auto w = [&](double n) {
        double s = 1.0;
        for (int m = 1; m < nbar; ++m)
            s += 2.0 * fm[m - 1] * std::cos(2.0 * kPi * m * (n - nf / 2.0 + 0.5) / nf);
        return s;
    };
    const double centre = w((nf - 1) / 2.0);
    for (int m = 0; m < ns; ++m) out[m] = static_cast<float>(w(m) / centre);
}

struct PosKey {
    uint32_t x, y, z;
    bool operator==(const PosKey &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct PosKeyHash {
    size_t operator()(const PosKey &k) const
    {
        uint64_t h = k.x * 0x9E3779B97F4A7C15ULL;
        h ^= (k.y + 0x7F4A7C159E3779B9ULL) + (h << 6) + (h >> 2);
        h ^= (k.z + 0x94D049BB133111EBULL) + (h << 6) + (h >> 2);
        return static_cast<size_t>(h);
    }
};
PosKey key_of(float x, float y, float z)
{
    // +0.0 and -0.0 compare equal, so normalise the sign of zero
    auto bits = [](float v) {
        if (v == 0.0F) v = 0.0F;
        uint32_t b;
        std::memcpy(&b, &v, 4);
        return b;
    };
    return PosKey{bits(x), bits(y), bits(z)};
}


} // namespace

struct tdc_processor {
    int device = 0;
    hipStream_t stream = nullptr;
    // timing events: ev[0] spans a whole run (and a streamed start); the
    // rest, in slots 1 and 2, bracket each correlate-sum launch of tdc_run,
    // whose iterations are enqueued back to back and read after one
    // synchronisation
    std::vector<std::array<hipEvent_t, 4>> ev;

    // FP32 drop-in inputs (borrowed)
    const float *raw = nullptr;
    const float *pri_times = nullptr;
    const float *sample_times = nullptr;
    const float *pos = nullptr;
    const float *vel = nullptr;
    const float *att = nullptr;
    int npri = 0, ns = 0;
    float mod_rate = 0.0F, start_freq = 0.0F;
    float bore[3] = {0.0F, 0.0F, 0.0F};
    float lambda_c = 0.0F;
    bool have_raw32 = false;

    // FP64 inputs of the precision form (borrowed)
    const double *sample_times64 = nullptr;
    const double *cpos64 = nullptr;
    double mod_rate64 = 0.0, start_freq64 = 0.0;
    bool have_raw64 = false;

    // focus grid (borrowed)
    const float *grid = nullptr;
    const double *grid64 = nullptr;
    int rows = 0, cols = 0;

    // options
    int form = TDC_FORM_PLAIN;
    int storage = TDC_RAW_C64;
    int engine = TDC_ENGINE_AUTO;
    int win_kind = TDC_WIN_HAMMING;
    double win_p0 = 25.0 / 46.0, win_p1 = 0.0; // tuning.h:31
    int force_ppl = 0, force_split = 0;
    int profile_taps = 0, profile_over = 0; // range-profile kernel (0, 0: automatic)
    bool profile_exact = false;             // range-profile engine: FP64 range for every chirp
    uint64_t budget = 0; // device bytes for one image's buffers (0: no limit)

    // device buffers
    DevBuf d_raw, d_pos, d_vel, d_att, d_kt, d_wr, d_rec, d_grid, d_part, d_img, d_cpos,
        d_kanchor, d_counters, d_atab, d_ascale, d_chirp, d_ptab, d_pcorr, d_ptw, d_pcpos;
    // the two buffer sets of chirp slabs streamed through the device, and the
    // FP64 image the slabs add into
    DevBuf r_raw[2], r_rec[2], r_cp[2], r_tab[2], r_scale[2], d_acc;
    hipEvent_t r_done[2] = {nullptr, nullptr};
    // host-to-device copies of the inputs (declared after the buffers it
    // fills, so it is destroyed, and its copies drained, before they are freed)
    tdc::Ingest ingest;

    // prepared state
    bool prepared = false;
    int run_form = TDC_FORM_PLAIN;
    int run_storage = TDC_RAW_C64;
    int run_engine = 0;
    bool doppler = false;
    bool chirp_path = false; // plain form: range (and window) once per (pixel, chirp)
    bool series = false;     // plain form, range per term: expansion of the range offset
    int ant_stride = 0;      // samples between the antenna states of consecutive chirps on the device
    // tdc_start uploads the focus grid behind the chirp inputs, in slabs that
    // overlap the correlate-sum of the slabs before them
    bool defer_grid = false;
    const void *grid_src = nullptr; // host grid in the device layout (FP32 or FP64 x, y, z)
    size_t grid_elem = 0;
    std::vector<double> grid64_host; // FP64 copy of an FP32 grid for the precision form
    tdc::Geom geom{};
    tdc::DopParams dp{};
    tdc::ProfileParams pp{};
    // Doppler staging parameters: 2 / lambda_c, centroid or fixed centre
    float st_tol = 0.0F, st_fc = 0.0F;
    int st_centroid = 0;
    int ppl = 1;
    double dk = 0.0;
    int64_t trivial_processed = 0;

    PinnedBuf image; // interleaved (re, im) of the last fetched image
    size_t image_px = 0;
    bool has_image = false;
    tdc_stats stats{};
};

namespace {

int ensure_device(tdc_processor *h)
{
    TDC_HIP(hipSetDevice(h->device));
    return TDC_OK;
}

// Power-of-two chirp split so that the launch has enough workgroups to fill
// 256 CUs several times over, while each split keeps a useful chirp count.
int choose_split(int ntiles, int npri, int force)
{
    if (force > 0) return force;
    // measured on 1024 chirps x 512 samples x 512^2 pixels: ~16k workgroups of 4 waves (8 rounds of 2048
    // resident ones) beat one resident round by 10-12%
    const int target_wgs = 16384;
    const int min_cps = 8;
    int s = 1;
    while (s < 64 && ntiles * s < target_wgs && (npri + 2 * s - 1) / (2 * s) >= min_cps) s *= 2;
    return s;
}

void make_geom(tdc::Geom &g, int npx, int npri, int ns, int ppl, int nsplit)
{
    g.npx = npx;
    g.npri = npri;
    g.ns = ns;
    g.ntiles = (npx + tdc::kBlock * ppl - 1) / (tdc::kBlock * ppl);
    g.nsplit = nsplit;
    g.cps = (npri + nsplit - 1) / nsplit;
    g.pstride = npx;
    const bool pow2 = (nsplit & (nsplit - 1)) == 0;
    g.xcd_map = pow2 ? 1 : 0;
    if (g.xcd_map && nsplit < 8) {
        const int per = 8 / nsplit;
        g.tiles_padded = (g.ntiles + per - 1) / per * per;
    } else {
        g.tiles_padded = g.ntiles;
    }
}


// True when every term's Doppler window value is exactly 1 and no (pixel,
// chunk) can be skipped, so the per-term window and the skip test can be
// left out without changing any result. Uses |f_D| <= (2/lambda)|v| and
// |f_c| <= (2/lambda)|v| |q|^2 |b| (a rotation by a non-unit quaternion
// scales by |q|^2), and excludes pixels that coincide with an antenna
// position (0/0 there makes the reference's window NaN, hence 0).
bool doppler_window_is_one(const tdc_processor *h, const tdc::AntennaScan &scan, float center,
                           float bw, bool centroid, bool taper)
{
    if (!(bw > 0.0F) || !std::isfinite(center)) return false;
    const size_t nterm = static_cast<size_t>(h->npri) * h->ns;
    const size_t npx = static_cast<size_t>(h->rows) * h->cols;
    if (!scan.has_vel || !scan.has_att || !h->pos || !scan.finite) return false;
    const double vmax2 = scan.vmax2, qmax2 = scan.qmax2;
    const double tol = 2.0 / h->lambda_c;
    const double vmax = std::sqrt(vmax2);
    const double bmag = std::sqrt(double(h->bore[0]) * h->bore[0] + double(h->bore[1]) * h->bore[1] +
                                  double(h->bore[2]) * h->bore[2]);
    const double fd_max = tol * vmax;
    const double fc_max = centroid ? tol * vmax * qmax2 * bmag : std::fabs(double(center));
    const double bound = (fd_max + fc_max) * (1.0 + 1e-4) + 1e-30;
    if (!std::isfinite(bound) || bound > double(bw) / 2.0) return false;
    // with the taper on, cos(2 pi d / bw - pi) must round to exactly -1
    if (taper && 2.0 * kPi * bound / double(bw) > 1e-5) return false;
    // a pixel at the origin meets the zero-filled probe rows; any other pixel
    // can only coincide with an antenna position inside the positions' bounding
    // box, so the set of positions is built only when such a pixel exists
    const float *lo = scan.lo, *hi = scan.hi;
    std::vector<size_t> inside;
    bool grid_ok = true;
    std::mutex mu;
    tdc::parallel_for(npx, 4096, [&](size_t i0, size_t i1) {
        std::vector<size_t> in;
        bool good = true;
        for (size_t i = i0; i < i1; ++i) {
            const float *g = h->grid + 3 * i;
            if (!std::isfinite(g[0]) || !std::isfinite(g[1]) || !std::isfinite(g[2]) ||
                (g[0] == 0.0F && g[1] == 0.0F && g[2] == 0.0F)) {
                good = false;
                break;
            }
            if (g[0] >= lo[0] && g[0] <= hi[0] && g[1] >= lo[1] && g[1] <= hi[1] &&
                g[2] >= lo[2] && g[2] <= hi[2])
                in.push_back(i);
        }
        std::lock_guard<std::mutex> lk(mu);
        grid_ok = grid_ok && good;
        inside.insert(inside.end(), in.begin(), in.end());
    });
    if (!grid_ok) return false;
    if (inside.empty()) return true;
    std::unordered_set<PosKey, PosKeyHash> seen;
    const float *last = nullptr;
    for (size_t e = 0; e < nterm; ++e) {
        const float *p = h->pos + 3 * e;
        if (last && p[0] == last[0] && p[1] == last[1] && p[2] == last[2]) continue;
        seen.insert(key_of(p[0], p[1], p[2]));
        last = p;
    }
    for (size_t i : inside) {
        const float *g = h->grid + 3 * i;
        if (seen.count(key_of(g[0], g[1], g[2]))) return false;
    }
    return true;
}

// True when the range offset of a term from its chirp's first position, dR,
// may come from the two-term expansion x - x^2 / (2 R0) (plain_tile): its
// remainder is at most k |delta|^3 / (2 R0^2) cycles of phase, with |delta|
// the antenna's largest offset within a chirp (the scan) and R0 at least the
// distance between the bounding boxes of the positions and the grid; the
// expansion is taken when that bound is below 5e-8 cycles and |delta| below
// 1% of the distance.
bool series_offset_ok(const tdc_processor *h, const tdc::AntennaScan &scan,
                      const std::vector<double> &k)
{
    if (!scan.finite || !h->grid) return false;
    const size_t npx = static_cast<size_t>(h->rows) * h->cols;
    if (npx == 0) return false;
    float glo[3], ghi[3];
    for (int a = 0; a < 3; ++a) glo[a] = ghi[a] = h->grid[a];
    std::mutex mu;
    bool finite = true;
    tdc::parallel_for(npx, 8192, [&](size_t i0, size_t i1) {
        float lo[3] = {glo[0], glo[1], glo[2]}, hi[3] = {ghi[0], ghi[1], ghi[2]};
        bool fin = true;
        for (size_t i = i0; i < i1; ++i) {
            for (int a = 0; a < 3; ++a) {
                const float v = h->grid[3 * i + a];
                fin = fin && std::isfinite(v);
                lo[a] = std::min(lo[a], v);
                hi[a] = std::max(hi[a], v);
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        finite = finite && fin;
        for (int a = 0; a < 3; ++a) {
            glo[a] = std::min(glo[a], lo[a]);
            ghi[a] = std::max(ghi[a], hi[a]);
        }
    });
    if (!finite) return false;
    double gap2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double g = std::max({0.0, double(glo[a]) - double(scan.hi[a]),
                                   double(scan.lo[a]) - double(ghi[a])});
        gap2 += g * g;
    }
    double kmax = 0.0;
    for (double v : k) kmax = std::max(kmax, std::fabs(v));
    const double delta = std::sqrt(scan.dmax2);
    return gap2 > 0.0 && delta * delta <= 1e-4 * gap2 &&
           kmax * delta * delta * delta <= 1e-7 * gap2;
}

// Queues a copy of `count` host values into `b` on the processor's ingest;
// `src` may be released on return. The compute stream waits for it after
// the next ingest fence.
template <class T> int upload(tdc_processor *h, DevBuf &b, const T *src, size_t count)
{
    TDC_HIP(b.ensure(sizeof(T) * count));
    if (count) TDC_HIP(h->ingest.put(b.p, src, sizeof(T) * count));
    return TDC_OK;
}

// The focus grid (x, y, z per pixel, in the form's precision) on the device:
// queued on the ingest now, or, for tdc_start, sized and left to the
// streamed pass.
template <class T> int place_grid(tdc_processor *h, const T *grid, int npx)
{
    const size_t count = 3 * static_cast<size_t>(npx);
    if (!h->defer_grid) return upload(h, h->d_grid, grid, count);
    TDC_HIP(h->d_grid.ensure(sizeof(T) * count));
    h->grid_src = grid;
    h->grid_elem = sizeof(T);
    return TDC_OK;
}

// Kernels that derive the device records from the raw samples and antenna
// states (window folding, storage conversion, per-chirp spectra or MFMA
// operands): part of every pass over a batch of input, so each tdc_run
// iteration launches them ahead of the correlate-sum.
int stage_inputs(tdc_processor *h)
{
    const int npri = h->npri, ns = h->ns;
    hipStream_t s = h->stream;
    if (h->run_form == TDC_FORM_PLAIN) {
        if (h->chirp_path) {
            TDC_HIP(tdc::launch_stage_terms(h->d_raw.as<float>(), h->d_kt.as<float>(),
                                            h->d_wr.as<float>(), npri, ns, h->d_rec.as<float4>(),
                                            s));
            float4 *cpos = h->d_chirp.as<float4>();
            if (h->doppler) {
                TDC_HIP(tdc::launch_stage_chirps(h->d_pos.as<float>(), h->d_vel.as<float>(),
                                                 h->d_att.as<float>(), npri, h->ant_stride, h->st_tol,
                                                 h->bore, h->st_centroid, h->st_fc, cpos,
                                                 cpos + npri, s));
            } else {
                TDC_HIP(tdc::launch_stage_chirps(h->d_pos.as<float>(), nullptr, nullptr, npri,
                                                 h->ant_stride,
                                                 0.0F, nullptr, 0, 0.0F, cpos, nullptr, s));
            }
        } else if (h->doppler) {
            TDC_HIP(tdc::launch_stage_doppler(h->d_raw.as<float>(), h->d_pos.as<float>(),
                                              h->d_vel.as<float>(), h->d_att.as<float>(),
                                              h->d_kt.as<float>(), h->d_wr.as<float>(), npri, ns,
                                              h->st_tol, h->bore, h->st_centroid, h->st_fc,
                                              h->d_rec.as<tdc::DopRec>(), s));
        } else {
            TDC_HIP(tdc::launch_stage_plain(h->d_raw.as<float>(), h->d_pos.as<float>(),
                                            h->d_kt.as<float>(), h->d_wr.as<float>(), npri, ns,
                                            h->d_rec.as<tdc::PlainRec>(), s));
            // the position of each chirp's first sample, the origin of the records' offsets
            TDC_HIP(tdc::launch_stage_chirps(h->d_pos.as<float>(), nullptr, nullptr, npri,
                                             h->ant_stride, 0.0F, nullptr, 0, 0.0F,
                                             h->d_chirp.as<float4>(), nullptr, s));
        }
        return TDC_OK;
    }
    if (h->run_engine == TDC_ENGINE_PROFILE) {
        // window and storage rounding folded into the per-chirp transform
        TDC_HIP(tdc::launch_stage_profile_folded(
            h->run_storage, h->d_raw.as<float>(), h->d_wr.as<float>(), h->d_pcorr.as<float>(),
            h->d_ptw.as<float2>(), h->d_pcpos.as<tdc::ChirpPos>(), npri, ns, h->pp.M, h->pp.Mp,
            h->pp.taps, h->pp.reverse, h->d_ptab.as<float2>(), s));
        return TDC_OK;
    }
    TDC_HIP(tdc::launch_stage_precision(h->run_storage, h->d_raw.as<float>(), h->d_wr.as<float>(),
                                        npri, ns, h->d_rec.p, s));
    if (h->run_engine == TDC_ENGINE_MATRIX) {
        TDC_HIP(tdc::launch_stage_matrix(h->run_storage, h->d_rec.p, npri, ns,
                                         tdc::matrix_parts(h->run_storage), h->d_atab.p,
                                         h->d_ascale.as<float>(), s));
    }
    return TDC_OK;
}

int prepare_plain(tdc_processor *h, float dop_center, float dop_bw, bool centroid, bool apply_rw,
                  bool taper)
{
    if (!h->have_raw32)
        return fail(TDC_ERR_STATE, "the plain form needs setRawData (FP32 inputs) before start()");
    if (!h->grid) return fail(TDC_ERR_STATE, "setFocusGrid must be called before start()");
    const int npri = h->npri, ns = h->ns, npx = h->rows * h->cols;
    const size_t nterm = static_cast<size_t>(npri) * ns;
    std::vector<float> wr, dk;
    std::vector<double> k(ns), kanchor;
    range_window(h->win_kind, h->win_p0, h->win_p1, ns, apply_rw, wr);
    for (int m = 0; m < ns; ++m) {
        const double f = double(h->start_freq) + double(h->mod_rate) * double(h->sample_times[m]);
        k[m] = 2.0 * f / kSpeedOfLight;
    }
    tdc::plain_phase_tables(k, dk, kanchor);
    tdc::AntennaScan scan;
    {
        tdc::Range checks("tdc_check_inputs");
        scan = tdc::scan_antenna(h->pos, h->vel, h->att, npri, ns);
        h->doppler = !doppler_window_is_one(h, scan, dop_center, dop_bw, centroid, taper);
    }
    if (h->doppler && (!h->vel || !h->att))
        return fail(TDC_ERR_INVALID_ARGUMENT, "the Doppler window needs velocity and attitude");
    // stop-and-hop input: with the position constant within every chirp the
    // range of a term depends only on its pixel and chirp; with velocity and
    // attitude constant too, so does its Doppler window
    h->chirp_path = scan.pos_const && (!h->doppler || (scan.vel_const && scan.att_const));
    h->series = !h->chirp_path && series_offset_ok(h, scan, k);
    hipStream_t s = h->stream;
    int rc;
    tdc::Range uploads("tdc_upload_inputs");
    TDC_HIP(h->ingest.begin(s));
    if ((rc = upload(h, h->d_raw, h->raw, nterm * 2)) != TDC_OK) return rc;
    if ((rc = upload(h, h->d_kt, dk.data(), ns)) != TDC_OK) return rc;
    if ((rc = upload(h, h->d_kanchor, kanchor.data(), kanchor.size())) != TDC_OK) return rc;
    if ((rc = upload(h, h->d_wr, wr.data(), ns)) != TDC_OK) return rc;
    if ((rc = place_grid(h, h->grid, npx)) != TDC_OK) return rc;
    if (h->chirp_path) {
        // one antenna state per chirp is all the per-chirp kernels read
        std::vector<float> cp(3 * static_cast<size_t>(npri)), cv, ca;
        if (h->doppler) {
            cv.resize(3 * static_cast<size_t>(npri));
            ca.resize(4 * static_cast<size_t>(npri));
        }
        for (size_t n = 0; n < static_cast<size_t>(npri); ++n) {
            const size_t e = n * ns;
            std::copy(h->pos + 3 * e, h->pos + 3 * e + 3, cp.begin() + 3 * n);
            if (h->doppler) {
                std::copy(h->vel + 3 * e, h->vel + 3 * e + 3, cv.begin() + 3 * n);
                std::copy(h->att + 4 * e, h->att + 4 * e + 4, ca.begin() + 4 * n);
            }
        }
        if ((rc = upload(h, h->d_pos, cp.data(), cp.size())) != TDC_OK) return rc;
        if ((rc = upload(h, h->d_vel, cv.data(), cv.size())) != TDC_OK) return rc;
        if ((rc = upload(h, h->d_att, ca.data(), ca.size())) != TDC_OK) return rc;
        h->ant_stride = 1;
    } else {
        if ((rc = upload(h, h->d_pos, h->pos, nterm * 3)) != TDC_OK) return rc;
        if (h->doppler) {
            if ((rc = upload(h, h->d_vel, h->vel, nterm * 3)) != TDC_OK) return rc;
            if ((rc = upload(h, h->d_att, h->att, nterm * 4)) != TDC_OK) return rc;
        }
        h->ant_stride = ns;
    }
    TDC_HIP(h->ingest.fence(s));
    if (h->doppler) {
        h->st_tol = 2.0F / h->lambda_c; // windowing.h:21, :33
        h->st_centroid = centroid ? 1 : 0;
        h->st_fc = dop_center;
        if (h->chirp_path) {
            TDC_HIP(h->d_rec.ensure(sizeof(float4) * nterm));
            TDC_HIP(h->d_chirp.ensure(sizeof(float4) * 2 * static_cast<size_t>(npri)));
        } else {
            TDC_HIP(h->d_rec.ensure(sizeof(tdc::DopRec) * nterm));
        }
        h->dp.half_bw = dop_bw / 2.0F;
        h->dp.inv_bw = 1.0F / dop_bw;
        h->dp.az_a = static_cast<float>(25.0 / 46.0); // tuning.h:38
        h->dp.taper = taper ? 1 : 0;
        h->dp.centroid = centroid ? 1 : 0;
        h->dp.fc_fixed = dop_center;
        TDC_HIP(h->d_counters.ensure(2 * sizeof(unsigned long long)));
    } else if (h->chirp_path) {
        TDC_HIP(h->d_rec.ensure(sizeof(float4) * nterm));
        TDC_HIP(h->d_chirp.ensure(sizeof(float4) * static_cast<size_t>(npri)));
    } else {
        TDC_HIP(h->d_rec.ensure(sizeof(tdc::PlainRec) * nterm));
        TDC_HIP(h->d_chirp.ensure(sizeof(float4) * static_cast<size_t>(npri)));
    }
    // measured best on 1024 chirps x 512 samples x 512^2 pixels: two pixels per
    // lane for the per-term plain records, one for the per-chirp range
    // (30.5 vs 31.4 ms at two) and for the Doppler kernels
    h->ppl = h->force_ppl ? h->force_ppl : ((h->doppler || h->chirp_path) ? 1 : 2);
    if (h->ppl > 4)
        return fail(TDC_ERR_INVALID_ARGUMENT, "the plain form runs 1, 2 or 4 pixels per lane");
    const int ntiles = (npx + tdc::kBlock * h->ppl - 1) / (tdc::kBlock * h->ppl);
    const int nsplit = choose_split(ntiles, npri, h->force_split);
    make_geom(h->geom, npx, npri, ns, h->ppl, nsplit);
    TDC_HIP(h->d_img.ensure(sizeof(float2) * static_cast<size_t>(npx)));
    if (nsplit > 1) TDC_HIP(h->d_part.ensure(sizeof(float2) * static_cast<size_t>(npx) * nsplit));
    h->run_form = TDC_FORM_PLAIN;
    h->run_engine = 0;
    h->run_storage = TDC_RAW_C64;
    return TDC_OK;
}

// The precision form's host side: input checks, the anchor and window
// tables, the chirp records, the engine and its parameters (h->dk, h->pp,
// h->ppl). Nothing touches the device.
struct PrecisionInputs {
    std::vector<double> kanchor;
    std::vector<tdc::ChirpPos> cp; // FP64 chirp positions
    std::vector<tdc::ChirpPos> pc; // range profile: positions from the scene origin, |.|^2 in pad
    std::vector<float> wr;
    const double *gridp = nullptr;
    int engine = 0;
    int plen = 0, mp = 0;
    std::shared_ptr<const tdc::ProfileTables> tabs;
};

int precision_inputs(tdc_processor *h, float dop_center, float dop_bw, bool centroid,
                     bool apply_rw, bool taper, PrecisionInputs &in)
{
    const int npri = h->npri, ns = h->ns, npx = h->rows * h->cols;
    std::vector<double> t(ns), cpos(static_cast<size_t>(npri) * 3);
    double f0, krate;
    if (h->have_raw64) {
        if (!h->grid64 && !h->grid)
            return fail(TDC_ERR_STATE, "setFocusGrid must be called before start()");
        // the FP64 API carries no velocity: the Doppler window must be identically 1
        if (taper || !(double(dop_bw) / 2.0 >= (centroid ? 0.0 : std::fabs(double(dop_center)))))
            return fail(TDC_ERR_UNSUPPORTED,
                        "the precision form applies no Doppler window; pass a band that "
                        "contains the window centre with the taper off");
        std::copy(h->sample_times64, h->sample_times64 + ns, t.begin());
        std::copy(h->cpos64, h->cpos64 + 3 * static_cast<size_t>(npri), cpos.begin());
        f0 = h->start_freq64;
        krate = h->mod_rate64;
    } else {
        if (!h->have_raw32) return fail(TDC_ERR_STATE, "setRawData must be called before start()");
        if (!h->grid) return fail(TDC_ERR_STATE, "setFocusGrid must be called before start()");
        tdc::Range checks("tdc_check_inputs");
        const tdc::AntennaScan scan = tdc::scan_antenna(h->pos, h->vel, h->att, npri, ns);
        if (!doppler_window_is_one(h, scan, dop_center, dop_bw, centroid, taper))
            return fail(TDC_ERR_UNSUPPORTED,
                        "the precision form needs a Doppler window that is identically 1 over "
                        "the data (use the plain form for Doppler windowing)");
        for (int n = 0; n < npri; ++n) {
            const float *p0 = h->pos + static_cast<size_t>(n) * ns * 3;
            // the scan compares bit patterns; positions that differ only in
            // the sign of zero are the same position here
            for (int m = 1; !scan.pos_const && m < ns; ++m) {
                const float *pm = p0 + 3 * static_cast<size_t>(m);
                if (pm[0] != p0[0] || pm[1] != p0[1] || pm[2] != p0[2])
                    return fail(TDC_ERR_UNSUPPORTED,
                                "the precision form needs one phase-centre position per chirp "
                                "(stop-and-hop); positions vary within chirp " +
                                    std::to_string(n));
            }
            cpos[3 * n] = p0[0];
            cpos[3 * n + 1] = p0[1];
            cpos[3 * n + 2] = p0[2];
        }
        for (int m = 0; m < ns; ++m) t[m] = h->sample_times[m];
        f0 = h->start_freq;
        krate = h->mod_rate;
    }
    if (h->grid64) {
        in.gridp = h->grid64;
    } else {
        std::vector<double> &grid64 = h->grid64_host;
        grid64.resize(static_cast<size_t>(npx) * 3);
        tdc::parallel_for(grid64.size(), size_t(1) << 16, [&](size_t i0, size_t i1) {
            for (size_t i = i0; i < i1; ++i) grid64[i] = h->grid[i];
        });
        in.gridp = grid64.data();
    }
    std::string why;
    if (!tdc::precision_tables(t.data(), ns, f0, krate, in.kanchor, h->dk, why))
        return fail(TDC_ERR_UNSUPPORTED, why);
    in.cp.resize(npri);
    for (int n = 0; n < npri; ++n)
        in.cp[n] = tdc::ChirpPos{cpos[3 * n], cpos[3 * n + 1], cpos[3 * n + 2], 0.0};
    range_window(h->win_kind, h->win_p0, h->win_p1, ns, apply_rw, in.wr);

    int engine = h->engine;
    in.plen = tdc::profile_length(ns);
    if (engine == TDC_ENGINE_AUTO) {
        if (h->force_ppl > 1) engine = TDC_ENGINE_VECTOR;
        else engine = in.plen <= tdc::kProfileMaxM ? TDC_ENGINE_PROFILE : TDC_ENGINE_MATRIX;
    }
    in.engine = engine;
    if (engine == TDC_ENGINE_PROFILE) {
        if (h->force_ppl > 1)
            return fail(TDC_ERR_INVALID_ARGUMENT, "the range-profile engine runs one pixel per lane");
        tdc::ProfileChoice pc;
        if (!tdc::profile_choice(ns, npx, h->profile_taps, h->profile_over, pc, why))
            return fail(TDC_ERR_UNSUPPORTED, why);
        in.plen = pc.M;
        in.tabs = tdc::profile_tables(ns, in.plen, pc.taps, pc.over);
        h->pp.taps = pc.taps;
        h->pp.over = pc.over;
        in.mp = in.plen + tdc::kProfilePad;
        // chirp positions from an origin inside the scene (the midpoint of the
        // first and last pixels), with |pos - o|^2 in the pad
        const double *gridp = in.gridp;
        const size_t last = 3 * (static_cast<size_t>(npx) - 1);
        double o[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 3; ++a) {
            o[a] = npx > 0 ? 0.5 * (gridp[a] + gridp[last + a]) : 0.0;
            if (!std::isfinite(o[a])) o[a] = 0.0;
        }
        in.pc.resize(npri);
        for (int n = 0; n < npri; ++n) {
            const tdc::ChirpPos &c = in.cp[n];
            const double x = c.x - o[0], y = c.y - o[1], z = c.z - o[2];
            in.pc[n] = tdc::ChirpPos{x, y, z, std::fma(z, z, std::fma(y, y, x * x))};
        }
        h->pp.ox = o[0];
        h->pp.oy = o[1];
        h->pp.oz = o[2];
        const double dt = ns > 1 ? (t[ns - 1] - t[0]) / (ns - 1) : 0.0;
        h->pp.dk = h->dk;
        h->pp.kh = 2.0 * (f0 + krate * (t[0] + (ns / 2) * dt)) / kSpeedOfLight;
        h->pp.M = in.plen;
        h->pp.Mp = in.mp;
        h->pp.dkm = std::fabs(h->dk) * in.plen;
        h->pp.reverse = h->dk < 0.0 ? 1 : 0;
        h->pp.incr = 0;
        if (!h->profile_exact && npx > 0) {
            double glo[3], ghi[3];
            tdc::grid_bounds(gridp, static_cast<size_t>(npx), glo, ghi);
            for (int a = 0; a < 3; ++a) {
                glo[a] -= o[a];
                ghi[a] -= o[a];
            }
            const double kmax = std::max(std::fabs(h->pp.kh), h->pp.dkm);
            h->pp.incr = tdc::profile_incremental_ok(in.pc.data(), npri, glo, ghi, kmax) ? 1 : 0;
        }
        h->ppl = 1;
    } else if (engine == TDC_ENGINE_MATRIX) {
        if (h->force_ppl > 1)
            return fail(TDC_ERR_INVALID_ARGUMENT, "the matrix engine runs one pixel per lane");
        h->ppl = 1;
    } else {
        h->ppl = h->force_ppl ? h->force_ppl : 2;
    }
    return TDC_OK;
}

// Device bytes of one chirp of the precision form (raw, staged raw but for
// the range-profile engine, which stages from the raw input, chirp record and
// the engine's per-chirp table) and of the buffers per pixel (grid, image and
// one partial image per split).
size_t precision_chirp_bytes(const tdc_processor *h, const PrecisionInputs &in)
{
    const size_t ns = static_cast<size_t>(h->ns);
    const size_t esz = h->storage == TDC_RAW_C64 ? 8 : 4;
    size_t b = 8 * ns + sizeof(tdc::ChirpPos);
    if (in.engine != TDC_ENGINE_PROFILE) b += esz * ns;
    if (in.engine == TDC_ENGINE_PROFILE)
        b += sizeof(float2) * static_cast<size_t>(in.mp) + sizeof(tdc::ChirpPos);
    else if (in.engine == TDC_ENGINE_MATRIX)
        b += tdc::matrix_table_bytes(1, h->ns, h->storage) + sizeof(float);
    return b;
}

// Bytes of one pixel's partial image per chirp split: FP32 from the
// range-profile engine, FP64 from the others.
size_t partial_bytes(int engine) { return engine == TDC_ENGINE_PROFILE ? sizeof(float2) : sizeof(double2); }
size_t precision_pixel_bytes(int nsplit, int engine)
{
    return 24 + 8 + partial_bytes(engine) * static_cast<size_t>(nsplit);
}

int prepare_precision(tdc_processor *h, float dop_center, float dop_bw, bool centroid,
                      bool apply_rw, bool taper, const PrecisionInputs *given)
{
    const int npri = h->npri, ns = h->ns, npx = h->rows * h->cols;
    const size_t nterm = static_cast<size_t>(npri) * ns;
    PrecisionInputs computed;
    int rc = TDC_OK;
    if (!given) {
        rc = precision_inputs(h, dop_center, dop_bw, centroid, apply_rw, taper, computed);
        if (rc != TDC_OK) return rc;
    }
    const PrecisionInputs &in = given ? *given : computed;
    const int ntiles = (npx + tdc::kBlock * h->ppl - 1) / (tdc::kBlock * h->ppl);
    const int nsplit = choose_split(ntiles, npri, h->force_split);
    if (h->budget > 0) {
        const double need = double(precision_chirp_bytes(h, in)) * npri +
                            double(precision_pixel_bytes(nsplit, in.engine)) * npx;
        if (need > double(h->budget))
            return fail(TDC_ERR_UNSUPPORTED,
                        "the image needs " + std::to_string(static_cast<long long>(need)) +
                            " device bytes, more than the device budget of " +
                            std::to_string(static_cast<unsigned long long>(h->budget)) +
                            ": tdc_start streams such an aperture in chirp slabs, tdc_prepare "
                            "keeps it resident");
    }
    hipStream_t s = h->stream;
    tdc::Range uploads("tdc_upload_inputs");
    TDC_HIP(h->ingest.begin(s));
    if ((rc = upload(h, h->d_raw, h->raw, nterm * 2)) != TDC_OK) return rc;
    if ((rc = upload(h, h->d_wr, in.wr.data(), ns)) != TDC_OK) return rc;
    if ((rc = upload(h, h->d_cpos, in.cp.data(), in.cp.size())) != TDC_OK) return rc;
    if ((rc = upload(h, h->d_kanchor, in.kanchor.data(), in.kanchor.size())) != TDC_OK) return rc;
    if ((rc = place_grid(h, in.gridp, npx)) != TDC_OK) return rc;
    const size_t esz = h->storage == TDC_RAW_C64 ? 8 : 4;
    // the range-profile engine stages straight from the raw input
    if (in.engine != TDC_ENGINE_PROFILE) TDC_HIP(h->d_rec.ensure(esz * nterm));
    h->doppler = false;
    if (in.engine == TDC_ENGINE_PROFILE) {
        if ((rc = upload(h, h->d_pcorr, in.tabs->corr.data(), in.tabs->corr.size())) != TDC_OK)
            return rc;
        if ((rc = upload(h, h->d_ptw, in.tabs->twiddle.data(), in.tabs->twiddle.size())) != TDC_OK)
            return rc;
        TDC_HIP(h->d_ptab.ensure(sizeof(float2) * static_cast<size_t>(npri) * in.mp));
        if ((rc = upload(h, h->d_pcpos, in.pc.data(), in.pc.size())) != TDC_OK) return rc;
    } else if (in.engine == TDC_ENGINE_MATRIX) {
        TDC_HIP(h->d_atab.ensure(tdc::matrix_table_bytes(npri, ns, h->storage)));
        TDC_HIP(h->d_ascale.ensure(sizeof(float) * static_cast<size_t>(npri)));
    }
    TDC_HIP(h->ingest.fence(s));
    h->run_engine = in.engine;
    make_geom(h->geom, npx, npri, ns, h->ppl, nsplit);
    TDC_HIP(h->d_img.ensure(sizeof(float2) * static_cast<size_t>(npx)));
    TDC_HIP(h->d_part.ensure(partial_bytes(in.engine) * static_cast<size_t>(npx) * nsplit));
    h->run_form = TDC_FORM_PRECISION;
    h->run_storage = h->storage;
    return TDC_OK;
}

} // namespace

namespace tdc {

int set_last_error(int code, const std::string &msg)
{
    t_err = msg;
    return code;
}

void clear_last_error() { t_err.clear(); }

// Anchor table and rotation step of the precision form. The recurrence
// advances by a constant phase step, so the sample times must be uniformly
// spaced (a few float32 ulps of the largest time are allowed, so that times
// given in float32 qualify; each anchor uses its own exact sample time).
namespace {
struct GaussLegendre64 {
    static constexpr int kNodes = 64;
    double node[kNodes], weight[kNodes];
    GaussLegendre64()
    {
        // Legendre nodes by Newton's method from the Chebyshev guesses
        for (int i = 0; i < kNodes; ++i) {
            double x = std::cos(kPi * (i + 0.75) / (kNodes + 0.5));
            double dp = 1.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = x;
                for (int k = 2; k <= kNodes; ++k) {
                    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = kNodes * (x * p1 - p0) / (x * x - 1.0);
                const double dx = p1 / dp;
                x -= dx;
                if (std::fabs(dx) < 1e-16) break;
            }
            node[i] = x;
            weight[i] = 2.0 / ((1.0 - x * x) * dp * dp);
        }
    }
};
} // namespace

std::shared_ptr<const ProfileTables> profile_tables(int ns, int M, int taps, int over)
{
    // computed once per (ns, M, taps, over) in the process: the tables depend on nothing else
    static std::mutex mu;
    static std::map<std::array<int, 4>, std::shared_ptr<const ProfileTables>> cache;
    const std::array<int, 4> key{ns, M, taps, over};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    static const GaussLegendre64 gl;
    constexpr int kNodes = GaussLegendre64::kNodes;
    const double half = taps / 2;
    const double b = profile_bfactor(taps, over) * kPi / half;
    // psi_hat is even in nu and the integrand even in u: sum the nodes of one
    // half twice, once per distance |i - h|
    double wexp[kNodes / 2], unode[kNodes / 2];
    for (int k = 0; k < kNodes / 2; ++k) {
        const double u = half * gl.node[k];
        unode[k] = u;
        wexp[k] = 2.0 * half * gl.weight[k] * std::exp(-b * u * u);
    }
    auto tab = std::make_shared<ProfileTables>();
    const int h = ns / 2;
    const int far = std::max(h, ns - 1 - h);
    std::vector<float> by_dist(far + 1);
    for (int d = 0; d <= far; ++d) {
        const double nu = double(d) / M;
        double v = 0.0;
        for (int k = 0; k < kNodes / 2; ++k) v += wexp[k] * std::cos(2.0 * kPi * nu * unode[k]);
        by_dist[d] = static_cast<float>(1.0 / v);
    }
    tab->corr.resize(ns);
    for (int i = 0; i < ns; ++i) tab->corr[i] = by_dist[std::abs(i - h)];
    tab->twiddle.resize(M / 2);
    for (int k = 0; k < M / 2; ++k) {
        const double a = 2.0 * kPi * k / M;
        tab->twiddle[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    std::lock_guard<std::mutex> lk(mu);
    return cache.emplace(key, std::move(tab)).first->second;
}

bool precision_tables(const double *t, int ns, double f0, double krate,
                      std::vector<double> &kanchor, double &dk, std::string &why)
{
    const double dt = ns > 1 ? (t[ns - 1] - t[0]) / (ns - 1) : 0.0;
    const double tmax = std::max(std::fabs(t[0]), std::fabs(t[ns - 1]));
    const double ttol = 1e-6 * std::fabs(dt) + 0x1p-21 * tmax + 1e-300;
    for (int m = 0; m < ns; ++m) {
        if (std::fabs(t[m] - (t[0] + m * dt)) > ttol) {
            why = "the precision form needs uniformly spaced sample times";
            return false;
        }
    }
    // one entry per anchor block, padded to whole block pairs; a padding
    // block holds no samples and takes its time from the uniform spacing
    const int nblk = tdc::anchor_entries(ns);
    kanchor.resize(nblk);
    for (int b = 0; b < nblk; ++b) {
        const int m = b * kAnchorBlock;
        const double tm = m < ns ? t[m] : t[0] + m * dt;
        kanchor[b] = 2.0 * (f0 + krate * tm) / kSpeedOfLight;
    }
    dk = 2.0 * krate * dt / kSpeedOfLight;
    return true;
}

void plain_phase_tables(const std::vector<double> &k, std::vector<float> &dk,
                        std::vector<double> &kanchor)
{
    const int ns = static_cast<int>(k.size());
    const int nblk = (ns + kAnchorBlock - 1) / kAnchorBlock;
    dk.resize(ns);
    kanchor.resize(nblk);
    for (int b = 0; b < nblk; ++b) {
        const int m0 = b * kAnchorBlock, m1 = std::min(ns, m0 + kAnchorBlock);
        double lo = k[m0], hi = k[m0];
        for (int m = m0; m < m1; ++m) {
            lo = std::min(lo, k[m]);
            hi = std::max(hi, k[m]);
        }
        const double kb = 0.5 * (lo + hi);
        kanchor[b] = std::isfinite(kb) ? kb : 0.0;
        for (int m = m0; m < m1; ++m) dk[m] = static_cast<float>(k[m] - kanchor[b]);
    }
}

} // namespace tdc

namespace tdc {

void make_launch_geom(Geom &g, int npx, int npri, int ns, int ppl, int nsplit)
{
    make_geom(g, npx, npri, ns, ppl, nsplit);
}

int noalloc_inputs(tdc_processor *h, float dop_center, float dop_bw, bool centroid, bool apply_rw,
                   bool taper, NoallocInputs &out)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (!h->have_raw32 && !h->have_raw64)
        return fail(TDC_ERR_STATE, "setRawData must be called before start()");
    if (!(h->form == TDC_FORM_PRECISION || h->have_raw64))
        return fail(TDC_ERR_UNSUPPORTED,
                    "the allocation-free entry point forms the precision form (FP64 inputs, or "
                    "tdc_set_form(TDC_FORM_PRECISION))");
    int rc = ensure_device(h);
    if (rc != TDC_OK) return rc;
    PrecisionInputs in;
    rc = precision_inputs(h, dop_center, dop_bw, centroid, apply_rw, taper, in);
    if (rc != TDC_OK) return rc;
    if (in.engine != TDC_ENGINE_PROFILE)
        return fail(TDC_ERR_UNSUPPORTED,
                    "the allocation-free entry point runs the range-profile engine (at most " +
                        std::to_string(kProfileMaxM / 2) + " samples per chirp)");
    out.device = h->device;
    out.stream = h->stream;
    out.npri = h->npri;
    out.ns = h->ns;
    out.npx = h->rows * h->cols;
    out.storage = h->storage;
    const int ntiles = (out.npx + kBlock - 1) / kBlock;
    out.nsplit = choose_split(ntiles, h->npri, h->force_split);
    out.forced = h->force_split > 0;
    out.raw = h->raw;
    out.grid = in.gridp;
    out.wr = std::move(in.wr);
    out.pc = std::move(in.pc);
    out.tabs = in.tabs;
    out.pp = h->pp;
    return TDC_OK;
}

void noalloc_finish(tdc_processor *h, const NoallocInputs &in, int nsplit, int workgroups,
                    double end_to_end_ms)
{
    h->prepared = false;
    tdc_stats st{};
    st.end_to_end_ms = end_to_end_ms;
    st.terms = static_cast<double>(in.npx) * in.npri * in.ns;
    st.form = TDC_FORM_PRECISION;
    st.raw_storage = in.storage;
    st.chirp_splits = nsplit;
    st.pixels_per_lane = 1;
    st.workgroups = workgroups;
    st.engine = TDC_ENGINE_PROFILE;
    st.profile_taps = in.pp.taps;
    st.profile_oversampling = in.pp.over;
    st.profile_incremental = in.pp.incr;
    h->stats = st;
    clear_last_error();
}

} // namespace tdc

extern "C" {

const char *tdc_last_error(void) { return t_err.c_str(); }

int tdc_abi_version(void) { return 20000; }

int tdc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tdc_create(int gpu_num, tdc_processor **out)
{
    if (!out) return fail(TDC_ERR_INVALID_ARGUMENT, "out handle pointer is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || gpu_num < 0 || gpu_num >= n) {
        return fail(TDC_ERR_DEVICE, std::string("Failed to set HIP device ") +
                                        std::to_string(gpu_num) + ": " +
                                        (e != hipSuccess ? hipGetErrorString(e)
                                                         : "no such device"));
    }
    auto *h = new tdc_processor();
    h->device = gpu_num;
    e = hipSetDevice(gpu_num);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        tdc_destroy(h);
        return fail(TDC_ERR_DEVICE, std::string("Failed to set HIP device: ") + hipGetErrorString(e));
    }
    *out = h;
    return ok();
}

int tdc_destroy(tdc_processor *h)
{
    if (!h) return ok();
    (void) hipSetDevice(h->device);
    if (h->stream) (void) hipStreamSynchronize(h->stream);
    for (auto &t : h->ev)
        for (hipEvent_t x : t)
            if (x) (void) hipEventDestroy(x);
    for (hipEvent_t x : h->r_done)
        if (x) (void) hipEventDestroy(x);
    if (h->stream) (void) hipStreamDestroy(h->stream);
    delete h;
    return ok();
}

int tdc_set_raw_data(tdc_processor *h, const float *raw_iq, const float *pri_times,
                     const float *sample_times, const float *position, const float *velocity,
                     const float *attitude, int n_pri, int n_samples, float mod_rate,
                     float start_freq, const float *body_boresight)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    // tdcprocessor.cpp:189-195
    if (n_pri <= 0) return fail(TDC_ERR_INVALID_ARGUMENT, "nPri must be greater than 0");
    if (n_samples <= 0) return fail(TDC_ERR_INVALID_ARGUMENT, "nSamples must be greater than 0");
    if (static_cast<int64_t>(n_pri) * n_samples > INT_MAX)
        return fail(TDC_ERR_INVALID_ARGUMENT, "nPri * nSamples must not exceed 2^31 - 1");
    if (!raw_iq || !sample_times || !position)
        return fail(TDC_ERR_INVALID_ARGUMENT, "rawData, sampleTimes and position must not be NULL");
    h->raw = raw_iq;
    h->pri_times = pri_times;
    h->sample_times = sample_times;
    h->pos = position;
    h->vel = velocity;
    h->att = attitude;
    h->npri = n_pri;
    h->ns = n_samples;
    h->mod_rate = mod_rate;
    h->start_freq = start_freq;
    for (int i = 0; i < 3; ++i) h->bore[i] = body_boresight ? body_boresight[i] : 0.0F;
    // tdcprocessor.cpp:204-208: centre wavelength in double, stored as float
    const double chirp_time = sample_times[n_samples - 1] - sample_times[0];
    const double end_freq = start_freq + static_cast<double>(mod_rate) * chirp_time;
    const double center_freq = (start_freq + end_freq) / 2.0;
    h->lambda_c = static_cast<float>(kSpeedOfLight / center_freq);
    h->have_raw32 = true;
    h->have_raw64 = false;
    h->prepared = false;
    return ok();
}

int tdc_set_focus_grid(tdc_processor *h, const float *grid, int n_rows, int n_cols)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n_rows < 0 || n_cols < 0)
        return fail(TDC_ERR_INVALID_ARGUMENT, "grid dimensions must not be negative");
    if (static_cast<int64_t>(n_rows) * n_cols > tdc::kMaxPixels)
        return fail(TDC_ERR_INVALID_ARGUMENT, "the focus grid must not exceed 2^30 pixels");
    if (!grid && n_rows > 0 && n_cols > 0) return fail(TDC_ERR_INVALID_ARGUMENT, "grid is NULL");
    h->grid = grid;
    h->grid64 = nullptr;
    h->rows = n_rows;
    h->cols = n_cols;
    h->prepared = false;
    return ok();
}

static int check_window(int kind, double p0, double p1)
{
    if (kind == TDC_WIN_HAMMING) {
        if (!(p0 >= 0.0 && p0 <= 1.0)) return fail(TDC_ERR_INVALID_ARGUMENT, "Hamming A must be in [0, 1]");
    } else if (kind == TDC_WIN_TAYLOR) {
        if (!(p0 >= 1.0 && p0 <= 64.0) || !(p1 > 0.0))
            return fail(TDC_ERR_INVALID_ARGUMENT, "Taylor needs 1 <= nbar <= 64 and sll > 0 dB");
    } else if (kind != TDC_WIN_HANN) {
        return fail(TDC_ERR_INVALID_ARGUMENT, "unknown window kind");
    }
    return TDC_OK;
}

int tdc_range_window(int kind, double p0, double p1, int n_samples, float *window)
{
    if (n_samples <= 0) return fail(TDC_ERR_INVALID_ARGUMENT, "nSamples must be greater than 0");
    if (!window) return fail(TDC_ERR_INVALID_ARGUMENT, "window is NULL");
    const int rc = check_window(kind, p0, p1);
    if (rc != TDC_OK) return rc;
    std::vector<float> w;
    range_window(kind, p0, p1, n_samples, true, w);
    std::copy(w.begin(), w.end(), window);
    return ok();
}

int tdc_set_range_window(tdc_processor *h, int kind, double p0, double p1)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    const int rc = check_window(kind, p0, p1);
    if (rc != TDC_OK) return rc;
    h->win_kind = kind;
    h->win_p0 = p0;
    h->win_p1 = p1;
    h->prepared = false;
    return ok();
}

int tdc_set_form(tdc_processor *h, int form)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (form != TDC_FORM_PLAIN && form != TDC_FORM_PRECISION)
        return fail(TDC_ERR_INVALID_ARGUMENT, "unknown form");
    h->form = form;
    h->prepared = false;
    return ok();
}

int tdc_set_raw_storage(tdc_processor *h, int storage)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (storage != TDC_RAW_C64 && storage != TDC_RAW_F16 && storage != TDC_RAW_BF16)
        return fail(TDC_ERR_INVALID_ARGUMENT, "unknown raw storage");
    h->storage = storage;
    h->prepared = false;
    return ok();
}

int tdc_set_engine(tdc_processor *h, int engine)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (engine != TDC_ENGINE_AUTO && engine != TDC_ENGINE_VECTOR && engine != TDC_ENGINE_MATRIX &&
        engine != TDC_ENGINE_PROFILE)
        return fail(TDC_ERR_INVALID_ARGUMENT, "unknown engine");
    h->engine = engine;
    h->prepared = false;
    return ok();
}

int tdc_set_raw_data_f64(tdc_processor *h, const float *raw_iq, const double *pri_times,
                         const double *sample_times, const double *chirp_position, int n_pri,
                         int n_samples, double mod_rate, double start_freq)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n_pri <= 0) return fail(TDC_ERR_INVALID_ARGUMENT, "nPri must be greater than 0");
    if (n_samples <= 0) return fail(TDC_ERR_INVALID_ARGUMENT, "nSamples must be greater than 0");
    if (static_cast<int64_t>(n_pri) * n_samples > INT_MAX)
        return fail(TDC_ERR_INVALID_ARGUMENT, "nPri * nSamples must not exceed 2^31 - 1");
    if (!raw_iq || !sample_times || !chirp_position)
        return fail(TDC_ERR_INVALID_ARGUMENT, "rawData, sampleTimes and position must not be NULL");
    (void) pri_times;
    h->raw = raw_iq;
    h->sample_times64 = sample_times;
    h->cpos64 = chirp_position;
    h->npri = n_pri;
    h->ns = n_samples;
    h->mod_rate64 = mod_rate;
    h->start_freq64 = start_freq;
    h->have_raw64 = true;
    h->have_raw32 = false;
    h->form = TDC_FORM_PRECISION;
    h->prepared = false;
    return ok();
}

int tdc_set_focus_grid_f64(tdc_processor *h, const double *grid, int n_rows, int n_cols)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (n_rows < 0 || n_cols < 0)
        return fail(TDC_ERR_INVALID_ARGUMENT, "grid dimensions must not be negative");
    if (static_cast<int64_t>(n_rows) * n_cols > tdc::kMaxPixels)
        return fail(TDC_ERR_INVALID_ARGUMENT, "the focus grid must not exceed 2^30 pixels");
    if (!grid && n_rows > 0 && n_cols > 0) return fail(TDC_ERR_INVALID_ARGUMENT, "grid is NULL");
    h->grid64 = grid;
    h->grid = nullptr;
    h->rows = n_rows;
    h->cols = n_cols;
    h->prepared = false;
    return ok();
}

int tdc_set_profile_kernel(tdc_processor *h, int taps, int oversampling)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (!((taps == 0 && oversampling == 0) || tdc::profile_kernel_valid(taps, oversampling)))
        return fail(TDC_ERR_INVALID_ARGUMENT,
                    "the range-profile kernels are (taps, oversampling) = (8, 2), (10, 2), "
                    "(12, 2), (6, 4) and (8, 4), or (0, 0) for the automatic choice");
    h->profile_taps = taps;
    h->profile_over = oversampling;
    h->prepared = false;
    return ok();
}

int tdc_set_profile_exact_range(tdc_processor *h, int exact)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    h->profile_exact = exact != 0;
    h->prepared = false;
    return ok();
}

int tdc_set_device_budget(tdc_processor *h, unsigned long long bytes)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    h->budget = bytes;
    h->prepared = false;
    return ok();
}

int tdc_set_launch(tdc_processor *h, int pixels_per_lane, int chirp_splits)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    if (pixels_per_lane != 0 && pixels_per_lane != 1 && pixels_per_lane != 2 &&
        pixels_per_lane != 4 && pixels_per_lane != 8)
        return fail(TDC_ERR_INVALID_ARGUMENT, "pixels per lane must be 0, 1, 2, 4 or 8");
    if (chirp_splits < 0 || chirp_splits > 4096)
        return fail(TDC_ERR_INVALID_ARGUMENT, "chirp splits must be in [0, 4096]");
    h->force_ppl = pixels_per_lane;
    h->force_split = chirp_splits;
    h->prepared = false;
    return ok();
}

namespace {
struct PrecisionInputs;
int prepare_inputs(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
                   int dop_centroid_win, int apply_range_win, int doppler_win_taper,
                   bool defer_grid, const PrecisionInputs *given = nullptr);
} // namespace

int tdc_prepare(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
                int dop_centroid_win, int apply_range_win, int doppler_win_taper)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    tdc::Range range("tdc_prepare");
    const int rc = prepare_inputs(h, doppler_win_center, doppler_bandwidth, dop_centroid_win,
                                  apply_range_win, doppler_win_taper, false);
    return rc == TDC_OK ? ok() : rc;
}

namespace {

// The correlate-sum kernel of the prepared form over the pixels
// [p0, p0 + g.npx): grid and partial images offset by p0, the partial
// images g.pstride apart (the image itself when the plain form has one split).
int launch_main(tdc_processor *h, const tdc::Geom &g, size_t p0, int *wgs)
{
    hipStream_t s = h->stream;
    if (h->run_form == TDC_FORM_PLAIN) {
        float2 *dst = (g.nsplit > 1 ? h->d_part.as<float2>() : h->d_img.as<float2>()) + p0;
        const float *grid = h->d_grid.as<float>() + 3 * p0;
        const float4 *cpos = h->d_chirp.as<float4>();
        const double *ka = h->d_kanchor.as<double>();
        if (h->doppler && h->chirp_path) {
            TDC_HIP(tdc::launch_doppler_chirp(h->ppl, h->d_rec.as<float4>(), cpos, cpos + g.npri,
                                              grid, ka, g, h->dp, dst,
                                              h->d_counters.as<unsigned long long>(), s, wgs));
        } else if (h->doppler) {
            TDC_HIP(tdc::launch_doppler(h->ppl, h->series, h->d_rec.as<tdc::DopRec>(), grid, ka, g,
                                        h->dp, dst, h->d_counters.as<unsigned long long>(), s,
                                        wgs));
        } else if (h->chirp_path) {
            TDC_HIP(tdc::launch_range(h->ppl, h->d_rec.as<float4>(), cpos, grid, ka, g, dst, s,
                                      wgs));
        } else {
            TDC_HIP(tdc::launch_plain(h->ppl, h->series, h->d_rec.as<tdc::PlainRec>(), cpos, grid,
                                      ka, g, dst, s, wgs));
        }
        return TDC_OK;
    }
    const double *grid = h->d_grid.as<double>() + 3 * p0;
    double2 *dst = h->d_part.as<double2>() + p0;
    if (h->run_engine == TDC_ENGINE_PROFILE) {
        TDC_HIP(tdc::launch_profile(h->d_ptab.as<float2>(), h->d_pcpos.as<tdc::ChirpPos>(), grid,
                                    h->pp, g, h->d_part.as<float2>() + p0, s, wgs));
    } else if (h->run_engine == TDC_ENGINE_MATRIX) {
        TDC_HIP(tdc::launch_precision_matrix(
            tdc::matrix_parts(h->run_storage), h->d_atab.p, h->d_ascale.as<float>(),
            h->d_cpos.as<tdc::ChirpPos>(), grid, h->d_kanchor.as<double>(), h->dk, g, dst, s,
            wgs));
    } else {
        TDC_HIP(tdc::launch_precision(h->ppl, h->run_storage, h->d_rec.p,
                                      h->d_cpos.as<tdc::ChirpPos>(), grid,
                                      h->d_kanchor.as<double>(), h->dk, g, dst, s, wgs));
    }
    return TDC_OK;
}

// The fixed-order sum of the splits' partial images into the image.
int launch_reduce(tdc_processor *h)
{
    const tdc::Geom &g = h->geom;
    if (h->run_form == TDC_FORM_PLAIN) {
        if (g.nsplit > 1)
            TDC_HIP(tdc::launch_finalize_f32(h->d_part.as<float2>(), g.nsplit, g.npx,
                                             h->d_img.as<float2>(), h->stream));
    } else if (h->run_engine == TDC_ENGINE_PROFILE) {
        TDC_HIP(tdc::launch_finalize_p32(h->d_part.as<float2>(), g.nsplit, g.npx,
                                         h->d_img.as<float2>(), h->stream));
    } else {
        TDC_HIP(tdc::launch_finalize_f64(h->d_part.as<double2>(), g.nsplit, g.npx,
                                         h->d_img.as<float2>(), h->stream));
    }
    return TDC_OK;
}

int ensure_events(tdc_processor *h, int count)
{
    while (static_cast<int>(h->ev.size()) < count) {
        std::array<hipEvent_t, 4> t{nullptr, nullptr, nullptr, nullptr};
        for (hipEvent_t &x : t) {
            const hipError_t e = hipEventCreate(&x);
            if (e != hipSuccess) {
                for (hipEvent_t y : t)
                    if (y) (void) hipEventDestroy(y);
                return fail(TDC_ERR_DEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
            }
        }
        h->ev.push_back(t);
    }
    return TDC_OK;
}

int finish_stats(tdc_processor *h, double main_ms, double total_ms, int wgs)
{
    const tdc::Geom &g = h->geom;
    tdc_stats &st = h->stats;
    st.main_kernel_ms = main_ms;
    st.kernel_ms = total_ms;
    st.terms = double(g.npx) * g.npri * g.ns;
    st.form = h->run_form;
    st.raw_storage = h->run_storage;
    st.doppler_active = h->doppler ? 1 : 0;
    st.chirp_splits = g.nsplit;
    st.pixels_per_lane = h->ppl;
    st.workgroups = wgs;
    st.engine = h->run_form == TDC_FORM_PRECISION ? h->run_engine : 0;
    st.profile_taps = st.engine == TDC_ENGINE_PROFILE ? h->pp.taps : 0;
    st.profile_oversampling = st.engine == TDC_ENGINE_PROFILE ? h->pp.over : 0;
    st.profile_incremental = st.engine == TDC_ENGINE_PROFILE ? h->pp.incr : 0;
    st.per_chirp = 0;
    if (h->run_form == TDC_FORM_PLAIN && h->chirp_path)
        st.per_chirp = TDC_PER_CHIRP_RANGE | (h->doppler ? TDC_PER_CHIRP_WINDOW : 0);
    if (h->run_form == TDC_FORM_PLAIN && !h->chirp_path && h->series)
        st.per_chirp = TDC_PER_TERM_SERIES;
    if (h->doppler && g.npx > 0) {
        unsigned long long c[2] = {0, 0};
        TDC_HIP(hipMemcpy(c, h->d_counters.p, sizeof c, hipMemcpyDeviceToHost));
        st.processed_chunks = static_cast<int64_t>(c[0]);
        st.skipped_chunks = static_cast<int64_t>(c[1]);
    } else {
        st.processed_chunks = h->trivial_processed;
        st.skipped_chunks = 0;
    }
    return TDC_OK;
}

// One pass of tdc_start with the focus grid still on the host: the staging
// kernels start as soon as the chirp inputs have landed, and the grid follows
// in two pixel slabs (the first a quarter of the tiles), each slab's
// correlate-sum launched as soon as its part of the grid has landed, so the
// larger part of the grid crosses PCIe while the first slab is formed. Every
// pixel is computed by the same tile code over the same splits as in
// tdc_run, so the image is the same bit for bit.
int run_streamed(tdc_processor *h)
{
    tdc::Range range("tdc_run");
    const tdc::Geom &g = h->geom;
    hipStream_t s = h->stream;
    int rc = ensure_events(h, 1);
    if (rc != TDC_OK) return rc;
    const std::array<hipEvent_t, 4> &evt = h->ev[0];
    const size_t tile = static_cast<size_t>(tdc::kBlock) * h->ppl;
    const size_t npx = static_cast<size_t>(g.npx);
    std::vector<size_t> cut{0};
    if (static_cast<int64_t>(g.ntiles) * g.nsplit >= 8192 && g.ntiles >= 8)
        cut.push_back(static_cast<size_t>(g.ntiles / 4) * tile);
    cut.push_back(npx);
    if (h->doppler) TDC_HIP(hipMemsetAsync(h->d_counters.p, 0, 2 * sizeof(unsigned long long), s));
    TDC_HIP(hipEventRecord(evt[0], s));
    if ((rc = stage_inputs(h)) != TDC_OK) return rc;
    TDC_HIP(hipEventRecord(evt[1], s));
    int wgs = 0;
    const char *src = static_cast<const char *>(h->grid_src);
    char *dst = static_cast<char *>(h->d_grid.p);
    const size_t px_bytes = 3 * h->grid_elem;
    for (size_t k = 0; k + 1 < cut.size(); ++k) {
        const size_t p0 = cut[k], p1 = cut[k + 1];
        if (p1 <= p0) continue;
        TDC_HIP(h->ingest.put(dst + p0 * px_bytes, src + p0 * px_bytes, (p1 - p0) * px_bytes));
        TDC_HIP(h->ingest.fence(s));
        tdc::Geom gs{};
        make_geom(gs, static_cast<int>(p1 - p0), g.npri, g.ns, h->ppl, g.nsplit);
        gs.pstride = g.npx;
        int w = 0;
        if ((rc = launch_main(h, gs, p0, &w)) != TDC_OK) return rc;
        wgs += w;
    }
    TDC_HIP(hipEventRecord(evt[2], s));
    if ((rc = launch_reduce(h)) != TDC_OK) return rc;
    TDC_HIP(hipEventRecord(evt[3], s));
    TDC_HIP(hipEventSynchronize(evt[3]));
    float a = 0.0F, b = 0.0F;
    TDC_HIP(hipEventElapsedTime(&a, evt[1], evt[2]));
    TDC_HIP(hipEventElapsedTime(&b, evt[0], evt[3]));
    return finish_stats(h, a, b, wgs);
}

// Slab plan of an aperture larger than the device budget: `sps` chirp splits
// of `cps` chirps per slab, `nsplit` splits in all, with cps = ceil(npri /
// nsplit) so that a resident pass with nsplit splits cuts the same splits.
// Each of the two buffer sets holds one slab; the grid, the image, the FP64
// sum and one partial image per split of a slab are held once.
struct RingPlan {
    int sps = 0, cps = 0, nsplit = 0;
};

int ring_plan(const tdc_processor *h, const PrecisionInputs &in, RingPlan &plan)
{
    const int npri = h->npri, npx = h->rows * h->cols;
    const double per_chirp = double(precision_chirp_bytes(h, in));
    const double tables = 8.0 * h->ns + 8.0 * in.kanchor.size() +
                          (in.tabs ? 4.0 * in.tabs->corr.size() + 8.0 * in.tabs->twiddle.size() : 0.0);
    const int ntiles = (npx + tdc::kBlock * h->ppl - 1) / (tdc::kBlock * h->ppl);
    const int nsplit_full = choose_split(ntiles, npri, 0);
    const int cps_full = (npri + nsplit_full - 1) / nsplit_full;
    for (int sps = 16; sps >= 1; sps /= 2) {
        // grid, image, FP64 sum and the slab's partial images
        const double fixed =
            (24.0 + 8.0 + 16.0 + double(partial_bytes(in.engine)) * sps) * npx + tables;
        const double chirps = (double(h->budget) - fixed) / (2.0 * per_chirp);
        int nsplit;
        if (h->force_split > 0) {
            nsplit = h->force_split;
        } else {
            const double cps_fit = std::floor(chirps / sps);
            if (cps_fit < 1.0) continue;
            const int cps_target = static_cast<int>(std::min<double>(cps_full, cps_fit));
            nsplit = (npri + cps_target - 1) / cps_target;
        }
        const int cps = (npri + nsplit - 1) / nsplit;
        if (double(cps) * sps > chirps) continue;
        plan.sps = sps;
        plan.cps = cps;
        plan.nsplit = nsplit;
        return TDC_OK;
    }
    return fail(TDC_ERR_UNSUPPORTED,
                "the device budget of " + std::to_string(static_cast<unsigned long long>(h->budget)) +
                    " bytes holds no slab of chirps of this image (two slabs of one split each, "
                    "with the grid, the image and its partial sums)");
}

// tdc_start of the precision form over an aperture larger than the device
// budget. The chirps pass through the device in slabs of whole chirp splits,
// alternating between two buffer sets: the copy stream fills one set while
// the compute stream stages and forms the slab held in the other, and waits
// for the set's previous slab to be formed before refilling it. Each slab's
// partial images are added in split order into an FP64 image that the last
// slab rounds into the image, so the result is bit for bit that of a
// resident pass with the same number of splits.
int run_ring(tdc_processor *h, const PrecisionInputs &in, const RingPlan &plan)
{
    tdc::Range range("tdc_run_slabs");
    const int npri = h->npri, ns = h->ns, npx = h->rows * h->cols;
    const size_t esz = h->storage == TDC_RAW_C64 ? 8 : 4;
    const int slab = plan.sps * plan.cps;
    hipStream_t s = h->stream;
    int rc = ensure_events(h, 1);
    if (rc != TDC_OK) return rc;
    TDC_HIP(hipStreamSynchronize(s));
    // the budget holds the slabs: release the buffers of a resident aperture
    for (DevBuf *b : {&h->d_raw, &h->d_rec, &h->d_cpos, &h->d_pcpos, &h->d_ptab, &h->d_atab,
                      &h->d_ascale, &h->d_part})
        b->release();
    for (int k = 0; k < 2; ++k) {
        TDC_HIP(h->r_raw[k].ensure(8 * static_cast<size_t>(ns) * slab));
        if (in.engine != TDC_ENGINE_PROFILE)
            TDC_HIP(h->r_rec[k].ensure(esz * static_cast<size_t>(ns) * slab));
        TDC_HIP(h->r_cp[k].ensure(sizeof(tdc::ChirpPos) * static_cast<size_t>(slab)));
        if (in.engine == TDC_ENGINE_PROFILE) {
            TDC_HIP(h->r_tab[k].ensure(sizeof(float2) * static_cast<size_t>(in.mp) * slab));
        } else if (in.engine == TDC_ENGINE_MATRIX) {
            TDC_HIP(h->r_tab[k].ensure(tdc::matrix_table_bytes(slab, ns, h->storage)));
            TDC_HIP(h->r_scale[k].ensure(sizeof(float) * static_cast<size_t>(slab)));
        }
        if (!h->r_done[k]) TDC_HIP(hipEventCreateWithFlags(&h->r_done[k], hipEventDisableTiming));
    }
    TDC_HIP(h->d_part.ensure(partial_bytes(in.engine) * static_cast<size_t>(npx) * plan.sps));
    TDC_HIP(h->d_acc.ensure(sizeof(double2) * static_cast<size_t>(npx)));
    TDC_HIP(h->d_img.ensure(sizeof(float2) * static_cast<size_t>(npx)));
    h->run_form = TDC_FORM_PRECISION;
    h->run_storage = h->storage;
    h->run_engine = in.engine;
    h->doppler = false;
    make_geom(h->geom, npx, npri, ns, h->ppl, plan.nsplit);

    const std::array<hipEvent_t, 4> &evt = h->ev[0];
    {
        tdc::Range uploads("tdc_upload_inputs");
        TDC_HIP(h->ingest.begin(s));
        if ((rc = upload(h, h->d_wr, in.wr.data(), ns)) != TDC_OK) return rc;
        if ((rc = upload(h, h->d_kanchor, in.kanchor.data(), in.kanchor.size())) != TDC_OK)
            return rc;
        if ((rc = upload(h, h->d_grid, in.gridp, 3 * static_cast<size_t>(npx))) != TDC_OK)
            return rc;
        if (in.engine == TDC_ENGINE_PROFILE) {
            if ((rc = upload(h, h->d_pcorr, in.tabs->corr.data(), in.tabs->corr.size())) != TDC_OK)
                return rc;
            if ((rc = upload(h, h->d_ptw, in.tabs->twiddle.data(), in.tabs->twiddle.size())) !=
                TDC_OK)
                return rc;
        }
    }
    TDC_HIP(hipEventRecord(evt[0], s));
    TDC_HIP(hipEventRecord(evt[1], s));
    const std::vector<tdc::ChirpPos> &recs = in.engine == TDC_ENGINE_PROFILE ? in.pc : in.cp;
    int wgs = 0;
    for (int k = 0, n0 = 0; n0 < npri; ++k, n0 += slab) {
        const int set = k & 1;
        const int nk = std::min(slab, npri - n0);
        if (k >= 2) TDC_HIP(h->ingest.wait(h->r_done[set]));
        TDC_HIP(h->ingest.put(h->r_raw[set].p, h->raw + 2 * static_cast<size_t>(n0) * ns,
                              8 * static_cast<size_t>(nk) * ns));
        TDC_HIP(h->ingest.put(h->r_cp[set].p, recs.data() + n0, sizeof(tdc::ChirpPos) * nk));
        TDC_HIP(h->ingest.fence(s));
        if (in.engine != TDC_ENGINE_PROFILE)
            TDC_HIP(tdc::launch_stage_precision(h->storage, h->r_raw[set].as<float>(),
                                                h->d_wr.as<float>(), nk, ns, h->r_rec[set].p, s));
        tdc::Geom gs{};
        make_geom(gs, npx, nk, ns, h->ppl, (nk + plan.cps - 1) / plan.cps);
        gs.cps = plan.cps;
        int w = 0;
        if (in.engine == TDC_ENGINE_PROFILE) {
            TDC_HIP(tdc::launch_stage_profile_folded(
                h->storage, h->r_raw[set].as<float>(), h->d_wr.as<float>(), h->d_pcorr.as<float>(),
                h->d_ptw.as<float2>(), h->r_cp[set].as<tdc::ChirpPos>(), nk, ns, in.plen, in.mp,
                h->pp.taps, h->pp.reverse, h->r_tab[set].as<float2>(), s));
            TDC_HIP(tdc::launch_profile(h->r_tab[set].as<float2>(),
                                        h->r_cp[set].as<tdc::ChirpPos>(), h->d_grid.as<double>(),
                                        h->pp, gs, h->d_part.as<float2>(), s, &w));
        } else if (in.engine == TDC_ENGINE_MATRIX) {
            const int parts = tdc::matrix_parts(h->storage);
            TDC_HIP(tdc::launch_stage_matrix(h->storage, h->r_rec[set].p, nk, ns, parts,
                                             h->r_tab[set].p, h->r_scale[set].as<float>(), s));
            TDC_HIP(tdc::launch_precision_matrix(
                parts, h->r_tab[set].p, h->r_scale[set].as<float>(),
                h->r_cp[set].as<tdc::ChirpPos>(), h->d_grid.as<double>(),
                h->d_kanchor.as<double>(), h->dk, gs, h->d_part.as<double2>(), s, &w));
        } else {
            TDC_HIP(tdc::launch_precision(h->ppl, h->storage, h->r_rec[set].p,
                                          h->r_cp[set].as<tdc::ChirpPos>(), h->d_grid.as<double>(),
                                          h->d_kanchor.as<double>(), h->dk, gs,
                                          h->d_part.as<double2>(), s, &w));
        }
        wgs += w;
        TDC_HIP(hipEventRecord(h->r_done[set], s));
        const bool last = n0 + slab >= npri;
        if (in.engine == TDC_ENGINE_PROFILE)
            TDC_HIP(tdc::launch_accumulate_p32(h->d_part.as<float2>(), gs.nsplit, npx,
                                               k == 0 ? 1 : 0, h->d_acc.as<double2>(),
                                               last ? h->d_img.as<float2>() : nullptr, s));
        else
            TDC_HIP(tdc::launch_accumulate_f64(h->d_part.as<double2>(), gs.nsplit, npx,
                                               k == 0 ? 1 : 0, h->d_acc.as<double2>(),
                                               last ? h->d_img.as<float2>() : nullptr, s));
    }
    TDC_HIP(hipEventRecord(evt[2], s));
    TDC_HIP(hipEventRecord(evt[3], s));
    TDC_HIP(hipEventSynchronize(evt[3]));
    float a = 0.0F, b = 0.0F;
    TDC_HIP(hipEventElapsedTime(&a, evt[1], evt[2]));
    TDC_HIP(hipEventElapsedTime(&b, evt[0], evt[3]));
    h->trivial_processed = static_cast<int64_t>(npx) * ((npri + tdc::kChunk - 1) / tdc::kChunk);
    return finish_stats(h, a, b, wgs);
}

int prepare_inputs(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
                   int dop_centroid_win, int apply_range_win, int doppler_win_taper,
                   bool defer_grid, const PrecisionInputs *given)
{
    if (!h->have_raw32 && !h->have_raw64)
        return fail(TDC_ERR_STATE, "setRawData must be called before start()");
    int rc = ensure_device(h);
    if (rc != TDC_OK) return rc;
    h->prepared = false;
    h->defer_grid = defer_grid;
    if (h->form == TDC_FORM_PRECISION || h->have_raw64) {
        rc = prepare_precision(h, doppler_win_center, doppler_bandwidth, dop_centroid_win != 0,
                               apply_range_win != 0, doppler_win_taper != 0, given);
    } else {
        rc = prepare_plain(h, doppler_win_center, doppler_bandwidth, dop_centroid_win != 0,
                           apply_range_win != 0, doppler_win_taper != 0);
    }
    if (rc != TDC_OK) return rc;
    const int nchunks = (h->npri + tdc::kChunk - 1) / tdc::kChunk;
    h->trivial_processed = static_cast<int64_t>(h->rows) * h->cols * nchunks;
    h->prepared = !defer_grid;
    return TDC_OK;
}

} // namespace

int tdc_run(tdc_processor *h, int iterations)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    tdc::Range range("tdc_run");
    if (!h->prepared) return fail(TDC_ERR_STATE, "tdc_prepare must succeed before tdc_run");
    if (iterations < 1) return fail(TDC_ERR_INVALID_ARGUMENT, "iterations must be >= 1");
    int rc = ensure_device(h);
    if (rc != TDC_OK) return rc;
    const tdc::Geom &g = h->geom;
    hipStream_t s = h->stream;
    double main_ms = 0.0;
    int wgs = 0;
    // events around each correlate-sum launch (a ring, read back every
    // kEventRing iterations) and one pair around the whole run
    constexpr int kEventRing = 64;
    if ((rc = ensure_events(h, std::min(iterations, kEventRing) + 1)) != TDC_OK) return rc;
    const std::array<hipEvent_t, 4> &span = h->ev[0];
    int pending = 0;
    auto drain = [&]() -> int {
        if (pending == 0) return TDC_OK;
        TDC_HIP(hipEventSynchronize(h->ev[pending][2]));
        for (int k = 1; k <= pending; ++k) {
            float a = 0.0F;
            TDC_HIP(hipEventElapsedTime(&a, h->ev[k][1], h->ev[k][2]));
            main_ms += a;
        }
        pending = 0;
        return TDC_OK;
    };
    const int ring = std::min(iterations, kEventRing);
    TDC_HIP(hipEventRecord(span[0], s));
    for (int it = 0; it < iterations; ++it) {
        if (g.npx == 0) break;
        const std::array<hipEvent_t, 4> &evt = h->ev[pending + 1];
        if (h->doppler) TDC_HIP(hipMemsetAsync(h->d_counters.p, 0, 2 * sizeof(unsigned long long), s));
        if ((rc = stage_inputs(h)) != TDC_OK) return rc;
        TDC_HIP(hipEventRecord(evt[1], s));
        if ((rc = launch_main(h, g, 0, &wgs)) != TDC_OK) return rc;
        TDC_HIP(hipEventRecord(evt[2], s));
        if ((rc = launch_reduce(h)) != TDC_OK) return rc;
        if (++pending == ring && it + 1 < iterations && (rc = drain()) != TDC_OK) return rc;
    }
    TDC_HIP(hipEventRecord(span[3], s));
    if ((rc = drain()) != TDC_OK) return rc;
    TDC_HIP(hipEventSynchronize(span[3]));
    float total = 0.0F;
    TDC_HIP(hipEventElapsedTime(&total, span[0], span[3]));
    if ((rc = finish_stats(h, main_ms / iterations, double(total) / iterations, wgs)) != TDC_OK)
        return rc;
    return ok();
}

namespace {

// The device image into the processor's page-locked image buffer (one DMA).
int fetch_image_now(tdc_processor *h)
{
    const size_t npx = static_cast<size_t>(h->rows) * h->cols;
    h->has_image = false;
    h->image_px = npx;
    if (npx) {
        TDC_HIP(h->image.ensure(npx * 2));
        TDC_HIP(hipMemcpyAsync(h->image.p, h->d_img.p, npx * sizeof(float2),
                               hipMemcpyDeviceToHost, h->stream));
        TDC_HIP(hipStreamSynchronize(h->stream));
    }
    h->has_image = true;
    return TDC_OK;
}

} // namespace

int tdc_fetch_image(tdc_processor *h)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    tdc::Range range("tdc_fetch_image");
    if (!h->prepared) return fail(TDC_ERR_STATE, "nothing has been run");
    int rc = ensure_device(h);
    if (rc != TDC_OK) return rc;
    if ((rc = fetch_image_now(h)) != TDC_OK) return rc;
    return ok();
}

int tdc_start(tdc_processor *h, float doppler_win_center, float doppler_bandwidth,
              int dop_centroid_win, int apply_range_win, int doppler_win_taper)
{
    if (!h) return fail(TDC_ERR_INVALID_ARGUMENT, "handle is NULL");
    tdc::Range range("tdc_start");
    const auto t0 = std::chrono::steady_clock::now();
    int rc = ensure_device(h);
    if (rc != TDC_OK) return rc;
    // with a device budget, a precision-form aperture that exceeds it streams
    // through the device in chirp slabs
    PrecisionInputs pin;
    bool have_pin = false;
    if (h->budget > 0 && (h->have_raw64 || (h->have_raw32 && h->form == TDC_FORM_PRECISION))) {
        tdc::Range prep("tdc_prepare");
        h->prepared = false;
        rc = precision_inputs(h, doppler_win_center, doppler_bandwidth, dop_centroid_win != 0,
                              apply_range_win != 0, doppler_win_taper != 0, pin);
        if (rc != TDC_OK) return rc;
        have_pin = true;
        const int npx = h->rows * h->cols;
        const int ntiles = (npx + tdc::kBlock * h->ppl - 1) / (tdc::kBlock * h->ppl);
        const int nsplit = choose_split(ntiles, h->npri, h->force_split);
        const double need = double(precision_chirp_bytes(h, pin)) * h->npri +
                            double(precision_pixel_bytes(nsplit, pin.engine)) * npx;
        if (need > double(h->budget) && npx > 0) {
            RingPlan plan;
            if ((rc = ring_plan(h, pin, plan)) != TDC_OK) return rc;
            if ((rc = run_ring(h, pin, plan)) != TDC_OK) return rc;
            if ((rc = fetch_image_now(h)) != TDC_OK) return rc;
            const auto t1 = std::chrono::steady_clock::now();
            h->stats.end_to_end_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            return ok();
        }
    }
    {
        tdc::Range prep("tdc_prepare");
        rc = prepare_inputs(h, doppler_win_center, doppler_bandwidth, dop_centroid_win,
                            apply_range_win, doppler_win_taper, true, have_pin ? &pin : nullptr);
    }
    if (rc == TDC_OK) {
        if (h->geom.npx > 0) {
            rc = run_streamed(h);
            if (rc == TDC_OK) h->prepared = true;
        } else {
            h->prepared = true;
            rc = tdc_run(h, 1);
        }
    }
    if (rc == TDC_OK) rc = tdc_fetch_image(h);
    if (rc != TDC_OK) return rc;
    const auto t1 = std::chrono::steady_clock::now();
    h->stats.end_to_end_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return ok();
}

int tdc_image_buffer(const tdc_processor *h, const float **image_iq)
{
    if (!h || !image_iq) return fail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    static const float empty_image[2] = {0.0F, 0.0F};
    if (!h->has_image) *image_iq = nullptr;
    else *image_iq = h->image_px == 0 ? empty_image : h->image.p;
    return ok();
}

int tdc_image_shape(const tdc_processor *h, int *n_rows, int *n_cols)
{
    if (!h || !n_rows || !n_cols) return fail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    *n_rows = h->rows;
    *n_cols = h->cols;
    return ok();
}

int tdc_get_stats(const tdc_processor *h, tdc_stats *out)
{
    if (!h || !out) return fail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = h->stats;
    return ok();
}

} // extern "C"
