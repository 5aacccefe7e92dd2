// tdc_internal.h -- shared between the HIP kernels and the host C ABI.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

struct tdc_processor;

namespace tdc {

// Chirps per skip-test chunk (reference tdc-lib/include/tuning.h:18).
constexpr int kChunk = 128;
// Threads per workgroup: four wave64 wavefronts.
constexpr int kBlock = 256;
// Samples per FP64 re-anchor block of the precision form.
constexpr int kAnchorBlock = 64;
// Matrix engine of the precision form: complex samples per contraction group
// (two k-slices of the 32x32x16 FP16 MFMA, 8 complex samples each), samples
// per anchor block, and samples per set of A tiles (two anchor blocks: the
// rows of one MFMA output).
constexpr int kGroup = 16;
constexpr int kSlice = 8;
constexpr int kMatrixBlock = 128;
constexpr int kQuadSamples = 2 * kMatrixBlock;
// Bytes of one A tile: 64 lanes x 8 FP16 values.
constexpr int kTileBytes = 64 * 16;


// Largest focus grid the setters accept (pixel indices, and the padded tile
// counts above them, stay inside 32-bit arithmetic).
constexpr int64_t kMaxPixels = int64_t(1) << 30;

// One (chirp, sample) term of the plain form with the antenna moving within
// the chirp, staged once per job and read by the workgroup through LDS.
struct alignas(32) PlainRec {
    float rr, ri; // raw sample times the range window
    float x, y, z; // antenna phase centre minus that of the chirp's first sample
    float k;      // dk_m = k_m - k_b, k_m = 2 (f0 + K t_m) / c in cycles per metre of range
                  // and k_b the anchor wavenumber of the sample's 64-sample block
    float pad0, pad1;
};

// Plain form with the per-term Doppler window (reference
// tdc-lib/src/tdckernels.cu:98-130): adds the velocity pre-scaled by
// 2 / lambda_c and the Doppler window centre of the term.
struct alignas(16) DopRec {
    float rr, ri;
    float x, y, z; // antenna phase centre (absolute: the skip test reads it)
    float k;       // dk_m, as in PlainRec
    float vx, vy, vz; // (2 / lambda_c) * velocity
    float fc;         // Doppler window centre (centroid or fixed), Hz
    float pad0, pad1;
};

// Records of the plain form when the antenna position is constant within
// every chirp (stop-and-hop): one float4 (rr, ri, dk_m, 0) per term, the raw
// sample times the range window and dk_m as in PlainRec, plus one float4 per
// chirp: its position (x, y, z, 0) and, with the Doppler window on, its
// velocity pre-scaled by 2 / lambda_c and window centre (vx, vy, vz, fc).
// Every plain kernel also reads the FP64 anchor wavenumbers k_b, one per
// kAnchorBlock samples (plain_phase_tables).

// Chirp phase centre of the precision form (FP64, padded to 32 bytes).
struct alignas(32) ChirpPos {
    double x, y, z, pad;
};

// Launch geometry shared by the correlate-sum kernels.
struct Geom {
    int npx;           // pixels
    int npri;          // chirps
    int ns;            // samples per chirp
    int ntiles;        // pixel tiles of kBlock * P pixels
    int nsplit;        // chirp ranges
    int cps;           // chirps per split
    int xcd_map;       // 1: XCD-aware block mapping (nsplit a power of two)
    int tiles_padded;  // tiles including the padding of the XCD mapping
    int pstride;       // distance between the partial images of consecutive splits (>= npx)
};

struct DopParams {
    float half_bw;   // dopplerBandwidth / 2
    float inv_bw;    // 1 / dopplerBandwidth
    float az_a;      // azimuth window parameter A (25/46)
    int taper;       // 1: raised-cosine taper, 0: rectangle
    int centroid;    // fc from the centroid (1) or fixed (0)
    float fc_fixed;  // dopplerWinCenter
};

// Range-profile engine of the precision form. With uniform sample times and
// one phase centre per chirp, the chirp's sum for a pixel at range R is
//   exp(j 2 pi k_h R) * sum_i c_i exp(j 2 pi (i - h) phi),  phi = dk R mod 1,
// the discrete-time Fourier transform of the windowed raw samples c_i
// (re-centred on sample h = ns / 2) at the frequency phi: the chirp's range
// profile. It is evaluated as a type-2 non-uniform FFT: once per chirp the
// samples, divided by the kernel's Fourier transform, go through an
// M-point FFT (M >= over * ns, a power of two, `over` the oversampling);
// each pixel then sums `taps` table entries around phi * M weighted by the
// Gaussian kernel exp(-b u^2), |u| <= taps / 2, b = k pi / (taps / 2) with k
// chosen per kernel for the smallest error. Built kernels (taps, over) and
// their error against the direct sum (rel-L2 of one chirp's profile at
// random frequencies, 512 Hann-windowed samples): (12, 2) 6.3e-8,
// (10, 2) 8.0e-7, (8, 2) 1.1e-5, (8, 4) 2.3e-6, (6, 4) 4.0e-5 -- fourfold
// oversampling buys the accuracy of about two more taps for twice the table.
// A Gaussian cut at |u| = taps / 2 cannot do much better than
// exp(-pi taps / 2) whatever the oversampling: (6, 8) measures 2.1e-5.
constexpr int kProfileTaps = 12; // the widest kernel
inline bool profile_kernel_valid(int taps, int over)
{
    return (over == 2 && (taps == 8 || taps == 10 || taps == 12)) ||
           (over == 4 && (taps == 6 || taps == 8));
}
// k of the kernel (taps, over); tdc_tiles.h holds the device constants of
// exp(-b u^2) generated from the same values
inline double profile_bfactor(int taps, int over)
{
    if (over == 4) return taps == 6 ? 0.87 : 0.88;
    return taps == 12 ? 0.79 : 0.80;
}
// The FFT of one chirp runs in the LDS, of which a gfx950 workgroup may
// address 160 KiB: radix-4 Stockham between two buffers (2 M entries) up to
// M = 8192, radix-2 in place in one buffer (M entries, 128 KiB) for M = 16384
constexpr int kProfileStockhamMaxM = 8192;
constexpr int kProfileMaxM = 16384;
// FFT length from which the per-chirp transform runs 1024 threads
constexpr int kProfileWideM = 4096;
// Entries after the M bins of a table row: wrap-around copies of its first
// bins (70), so that a wave's 64-entry window starting at any bin lies in the
// row, and the chirp's position increment (the last 2)
constexpr int kProfilePad = 72;
struct ProfileParams {
    double dk; // 2 K dt / c: cycles per sample per metre of range
    double kh; // 2 (f0 + K t_h) / c: wavenumber at the re-centring sample h
    double ox, oy, oz; // origin the chirp positions are given from (a pixel of the grid)
    int M;     // FFT length
    int Mp;    // table row length, M + kProfilePad
    double dkm; // |dk| M: the bin position of a pixel at range R is frac(dkm R / M) M
    int taps;   // kernel width
    int over;   // oversampling, M >= over * ns
    int reverse; // 1 for dk < 0 (down-chirps): table rows hold B[-m] instead of B[m]
    int incr;    // 1: within a block of chirps, the range, bin position and anchor advance
                 // by FP32 increments from the block's first chirp (profile_incremental_ok)
};
// The last two entries of a table row (never read as taps, which reach at
// most M + kProfileTaps - 2, nor as window entries, which reach at most
// M + 62) hold the chirp's position increment: float4
// (c_n - c_{n-1}, (|c_n|^2 - |c_{n-1}|^2) / 2) of the origin-relative positions
constexpr int kProfileDeltaEntry = kProfilePad - 2;
static_assert(kProfileDeltaEntry >= 63, "a 64-entry window from bin M - 1 must end before the increment");

// Whether the range-profile tile may advance a pixel's range by FP32
// increments within a block of chirps (the block's first chirp is exact in
// FP64) instead of evaluating it in FP64 for every chirp. A step changes the
// range by dR = delta / (R_n + R_{n-1}), delta = R_n^2 - R_{n-1}^2 =
// (|c_n|^2 - |c_{n-1}|^2) - 2 p . (c_n - c_{n-1}), formed in FP32 from
// rounded inputs by three FMAs with absolute error below eps32 (4 |c|max +
// 5 |p|max) |dc|max (taken as eps32 (5 |c|max + 6 |p|max) |dc|max), and taken from the
// series dR = e0 (1 - t (1 - 2 t)), e0 = delta / (2 R_{n-1}), t = e0 / (2
// R_{n-1}), whose remainder is below 0.625 e0^4 / R^3 with |e0| <= |dc|max
// (1 + |dc|max / R); the FP32 products that form dR (the reciprocal of R,
// e0 and the series) add at most 4 eps32 |e0|, and the FP32 roundings of k_h
// and |dk| M that advance the phase and bin by dR one more eps32 |e0|. Their
// sum, times the largest wavenumber kmax, must stay below kIncrementalCycles
// per step; R is bounded below by the distance between the
// bounding boxes of the chirp positions and the pixels. `c`: origin-relative
// chirp positions; `gmin`/`gmax`: the grid's bounding box relative to the
// same origin.
constexpr double kIncrementalCycles = 2e-6;
inline bool profile_incremental_ok(const ChirpPos *c, int npri, const double gmin[3],
                                   const double gmax[3], double kmax)
{
    if (npri < 2) return false;
    double cmin[3], cmax[3], step2 = 0.0, cnorm2 = 0.0;
    for (int a = 0; a < 3; ++a) cmin[a] = cmax[a] = (&c[0].x)[a];
    for (int n = 0; n < npri; ++n) {
        const double v[3] = {c[n].x, c[n].y, c[n].z};
        double s2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(v[a])) return false;
            cmin[a] = std::min(cmin[a], v[a]);
            cmax[a] = std::max(cmax[a], v[a]);
            if (n > 0) {
                const double d = v[a] - (&c[n - 1].x)[a];
                s2 += d * d;
            }
        }
        step2 = std::max(step2, s2);
        cnorm2 = std::max(cnorm2, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    double gap2 = 0.0, pnorm2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(gmin[a]) || !std::isfinite(gmax[a])) return false;
        const double gap = std::max({0.0, gmin[a] - cmax[a], cmin[a] - gmax[a]});
        gap2 += gap * gap;
        const double pa = std::max(std::fabs(gmin[a]), std::fabs(gmax[a]));
        pnorm2 += pa * pa;
    }
    const double rmin = std::sqrt(gap2), dc = std::sqrt(step2);
    if (!(rmin > 0.0) || !(kmax > 0.0)) return false;
    const double eps32 = 1.0 / 16777216.0;
    const double e0 = dc * (1.0 + dc / rmin);
    const double err_delta = eps32 * (5.0 * std::sqrt(cnorm2) + 6.0 * std::sqrt(pnorm2)) * dc;
    const double err_dr = err_delta / (2.0 * rmin) + 5.0 * eps32 * e0 +
                          0.625 * e0 * e0 * e0 * e0 / (rmin * rmin * rmin);
    return kmax * err_dr <= kIncrementalCycles;
}
inline int profile_length(int ns, int over = 2)
{
    int m = 16;
    while (m < over * static_cast<int64_t>(ns)) m *= 2;
    return m;
}
// The range-profile kernel for chirps of ns samples and an image of npx
// pixels: `taps` and `over` as requested, or (0, 0) for the automatic choice:
// 6 taps with fourfold oversampling when the transform fits (M <=
// kProfileMaxM) and the image has at least kProfileOver4Pixels pixels, else
// 10 taps with twofold oversampling. The fourfold table costs one more FFT
// pass and twice the points per chirp (measured on 1024 x 512 x 512^2:
// staging 10 -> 15 us); each two taps fewer save 8 of ~74 VALU instructions
// per (pixel, chirp) (the correlate-sum there: (10, 2) 0.460 ms, (8, 4)
// 0.410 ms, (6, 4) 0.372 ms), so the fourfold table pays from ~3e4 pixels
// per chirp. (6, 4) keeps the image within ~3e-5 of the direct sum, a third
// of the 1e-4 bound; (8, 4) and the twofold kernels are there for tighter
// bounds.
// False with a reason when the request is not a built kernel or its
// transform is too long.
constexpr int64_t kProfileOver4Pixels = 32768;
struct ProfileChoice {
    int taps = 0, over = 0, M = 0;
};
inline bool profile_choice(int ns, int64_t npx, int taps, int over, ProfileChoice &c,
                           std::string &why)
{
    if (taps == 0 && over == 0) {
        if (profile_length(ns, 4) <= kProfileMaxM && npx >= kProfileOver4Pixels) {
            taps = 6;
            over = 4;
        } else {
            taps = 10;
            over = 2;
        }
    }
    if (!profile_kernel_valid(taps, over)) {
        why = "the range-profile engine's kernels are (taps, oversampling) = (8, 2), (10, 2), "
              "(12, 2), (6, 4) and (8, 4)";
        return false;
    }
    const int M = profile_length(ns, over);
    if (M > kProfileMaxM) {
        why = "the range-profile engine takes at most " + std::to_string(kProfileMaxM / over) +
              " samples per chirp with " + std::to_string(over) + "-fold oversampling";
        return false;
    }
    c.taps = taps;
    c.over = over;
    c.M = M;
    return true;
}

// Kernel launch wrappers (defined in tdc_kernels.hip). All enqueue on `s`.
// `dk`: per-sample dk_m; `kanchor`: the FP64 anchor wavenumbers; `cpos`: the
// position of each chirp's first sample (launch_stage_chirps).
hipError_t launch_stage_plain(const float *raw, const float *pos, const float *dk, const float *wr,
                              int npri, int ns, PlainRec *out, hipStream_t s);
hipError_t launch_stage_doppler(const float *raw, const float *pos, const float *vel,
                                const float *att, const float *dk, const float *wr, int npri,
                                int ns, float two_over_lambda, const float *bore, int centroid,
                                float fc_fixed, DopRec *out, hipStream_t s);
// `series`: the range offset within a chirp from the two-term expansion
// (plain_tile; the host has bounded the antenna's motion within a chirp).
hipError_t launch_plain(int ppl, bool series, const PlainRec *rec, const float4 *cpos,
                        const float *grid, const double *kanchor, const Geom &g, float2 *out,
                        hipStream_t s, int *workgroups);
hipError_t launch_doppler(int ppl, bool series, const DopRec *rec, const float *grid,
                          const double *kanchor, const Geom &g, const DopParams &dp, float2 *out,
                          unsigned long long *counters, hipStream_t s, int *workgroups);
// Stop-and-hop records: terms (rr, ri, dk_m, 0) and per chirp the position of
// its first sample; with `vel` non-null also (2/lambda_c) v and the window
// centre of its first sample (the values stage_doppler writes for it).
hipError_t launch_stage_terms(const float *raw, const float *dk, const float *wr, int npri, int ns,
                              float4 *out, hipStream_t s);
hipError_t launch_stage_chirps(const float *pos, const float *vel, const float *att, int npri,
                               int ns, float two_over_lambda, const float *bore, int centroid,
                               float fc_fixed, float4 *cpos, float4 *cvel, hipStream_t s);
// Range evaluated once per (pixel, chirp): position constant within each chirp.
hipError_t launch_range(int ppl, const float4 *rec, const float4 *cpos, const float *grid,
                        const double *kanchor, const Geom &g, float2 *out, hipStream_t s,
                        int *workgroups);
// Range and Doppler window evaluated once per (pixel, chirp): position,
// velocity and attitude constant within each chirp.
hipError_t launch_doppler_chirp(int ppl, const float4 *rec, const float4 *cpos,
                                const float4 *cvel, const float *grid, const double *kanchor,
                                const Geom &g, const DopParams &dp, float2 *out,
                                unsigned long long *counters, hipStream_t s, int *workgroups);
hipError_t launch_finalize_f32(const float2 *partial, int nsplit, int npx, float2 *img,
                               hipStream_t s);

hipError_t launch_stage_precision(int storage, const float *raw, const float *wr, int npri, int ns,
                                  void *out, hipStream_t s);
hipError_t launch_precision(int ppl, int storage, const void *raw, const ChirpPos *cpos,
                            const double *grid, const double *kanchor, double dk, const Geom &g,
                            double2 *out, hipStream_t s, int *workgroups);
// Matrix engine: A tiles of every (chirp, block pair) from the staged raw
// samples (in `storage` format), `parts` = 2 (FP16 hi + lo) for c64 raw and
// 1 for raw that FP16 holds exactly (f16, bf16); one scale per chirp.
hipError_t launch_stage_matrix(int storage, const void *staged, int npri, int ns, int parts,
                               void *atab, float *inv_scale, hipStream_t s);
hipError_t launch_precision_matrix(int parts, const void *atab, const float *inv_scale,
                                   const ChirpPos *cpos, const double *grid,
                                   const double *kanchor, double dk, const Geom &g,
                                   double2 *out, hipStream_t s, int *workgroups);
__host__ __device__ inline int matrix_quads(int ns) { return (ns + kQuadSamples - 1) / kQuadSamples; }
inline int matrix_parts(int storage) { return storage == 0 ? 2 : 1; }
inline size_t matrix_table_bytes(int npri, int ns, int storage)
{
    return static_cast<size_t>(npri) * matrix_quads(ns) * 2 * matrix_parts(storage) * kTileBytes;
}
// Entries of the anchor table (one per 64 samples, padded to whole A tile
// sets so that the matrix engine's last anchor exists).
inline int anchor_entries(int ns) { return 4 * matrix_quads(ns); }
hipError_t launch_finalize_f64(const double2 *partial, int nsplit, int npx, float2 *img,
                               hipStream_t s);
// The same FP64 sum over FP32 partial images (the range-profile engine's).
hipError_t launch_finalize_p32(const float2 *partial, int nsplit, int npx, float2 *img,
                               hipStream_t s);
// acc = (first ? 0 : acc) + the nsplit partial images in split order; with img
// non-null the sum is rounded into img instead of stored in acc (_p32: FP32
// partial images).
hipError_t launch_accumulate_f64(const double2 *partial, int nsplit, int npx, int first,
                                 double2 *acc, float2 *img, hipStream_t s);
hipError_t launch_accumulate_p32(const float2 *partial, int nsplit, int npx, int first,
                                 double2 *acc, float2 *img, hipStream_t s);
// Range-profile tables [npri][Mp] from the staged raw samples (in `storage`
// format, range window folded in); corr[i] = 1 / psi_hat((i - h) / M);
// twiddle[k] = exp(+j 2 pi k / M), k < M / 2; rows rotated by taps / 2 - 1
// bins for a kernel of `taps` taps, and reversed (B[-m]) when `reverse`;
// Mp = M + kProfilePad, and each row ends with the chirp's position increment
// from `cpos` (the records launch_profile reads; kProfileDeltaEntry).
hipError_t launch_stage_profile(int storage, const void *staged, const float *corr,
                                const float2 *twiddle, const ChirpPos *cpos, int npri, int ns,
                                int M, int Mp, int taps, int reverse, float2 *table, hipStream_t s);
// The same tables from the complex64 input: each sample times the range
// window wr[i], rounded as `storage` stores it (bit-identical to staging it
// with launch_stage_precision first).
hipError_t launch_stage_profile_folded(int storage, const float *raw, const float *wr,
                                       const float *corr, const float2 *twiddle,
                                       const ChirpPos *cpos, int npri, int ns, int M, int Mp,
                                       int taps, int reverse, float2 *table, hipStream_t s);
// `cpos`: chirp positions minus the origin (pp.ox, oy, oz), with |pos - o|^2 in
// pad; `out`: one FP32 partial image per chirp split (sum them in FP64 with
// launch_finalize_p32 / launch_accumulate_p32).
hipError_t launch_profile(const float2 *table, const ChirpPos *cpos, const double *grid,
                          const ProfileParams &pp, const Geom &g, float2 *out, hipStream_t s,
                          int *workgroups);

// A named roctx range over a scope: host phases (prepare, run, fetch, job
// uploads, serving steps) appear beside the kernels in a profiler's marker
// trace; without a profiler attached the calls return at once.
class Range {
  public:
    explicit Range(const char *name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range &) = delete;
    Range &operator=(const Range &) = delete;
};

// Host helpers shared by the processor and the server (tdc_capi.cpp).
// The thread-local message returned by tdc_last_error().
int set_last_error(int code, const std::string &msg);
void clear_last_error();
// Host tables of the range-profile engine for ns samples, an M-point FFT and
// the kernel (taps, over): corr[i] = 1 / psi_hat((i - ns/2) / M), psi_hat the
// Fourier transform of the truncated Gaussian exp(-b u^2), |u| <= taps/2
// (64-point Gauss-Legendre quadrature in FP64); twiddle[k] = exp(+j 2 pi k /
// M), k < M / 2. Shared, computed once per (ns, M, taps, over).
struct ProfileTables {
    std::vector<float> corr;
    std::vector<float2> twiddle;
};
std::shared_ptr<const ProfileTables> profile_tables(int ns, int M, int taps, int over);
bool precision_tables(const double *t, int ns, double f0, double krate,
                      std::vector<double> &kanchor, double &dk, std::string &why);
// Phase tables of the plain form from the per-sample wavenumbers
// k[m] = 2 (f0 + K t_m) / c (FP64): per kAnchorBlock samples the anchor k_b,
// the midpoint of the block's smallest and largest k (FP64), and per sample
// dk[m] = k[m] - k_b rounded once to FP32.
void plain_phase_tables(const std::vector<double> &k, std::vector<float> &dk,
                        std::vector<double> &kanchor);

// Launch geometry of the one-job kernels (tdc_capi.cpp): pixels, chirps,
// samples, pixels per lane, chirp splits.
void make_launch_geom(Geom &g, int npx, int npri, int ns, int ppl, int nsplit);

// Host side of the allocation-free entry point (tdc_noalloc.hip): the
// precision form's range-profile inputs of a processor's current data, as
// tdc_start computes them (tdc_capi.cpp).
struct NoallocInputs {
    int device = 0;
    hipStream_t stream = nullptr;
    int npri = 0, ns = 0, npx = 0, storage = 0;
    int nsplit = 1;        // the chirp split count tdc_start chooses
    bool forced = false;   // that count was set with tdc_set_launch
    const float *raw = nullptr;   // host complex64 [npri][ns]
    const double *grid = nullptr; // host FP64 [npx][3]
    std::vector<float> wr;        // range window [ns]
    std::vector<ChirpPos> pc;     // chirp positions from the scene origin, |.|^2 in pad
    std::shared_ptr<const ProfileTables> tabs;
    ProfileParams pp{};
};
// Validates the processor's inputs for the allocation-free entry point and
// fills `in`; on failure returns the status and sets the message.
int noalloc_inputs(tdc_processor *h, float dop_center, float dop_bw, bool centroid, bool apply_rw,
                   bool taper, NoallocInputs &in);
// Records the statistics of an allocation-free pass; the processor's
// prepared state (which that pass did not use) must be prepared again.
void noalloc_finish(tdc_processor *h, const NoallocInputs &in, int nsplit, int workgroups,
                    double end_to_end_ms);

// Host ingest (tdc_host.cpp). Host threads for passes over one image's inputs.
constexpr unsigned kHostThreads = 8;
// Runs fn(begin, end) over contiguous parts of [0, n), each of at least
// `grain` items, on the host pool and the calling thread; returns when all
// parts have run.
void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)> &fn);
// Bounding box of npx FP64 points [npx][3] (on the host thread pool); NaN
// bounds when a coordinate is NaN.
void grid_bounds(const double *grid, size_t npx, double lo[3], double hi[3]);

// One pass over per-sample antenna states [npri][ns][3|3|4] (vel and att may
// be null): finiteness, the largest |v|^2 and |q|^2, the bounding box of the
// positions, the largest squared distance of a sample's position from its
// chirp's first, and whether each of pos, vel, att is bitwise the same for
// all samples of every chirp.
struct AntennaScan {
    bool has_vel = false, has_att = false;
    bool finite = true;
    bool pos_const = true, vel_const = true, att_const = true;
    double vmax2 = 0.0, qmax2 = 0.0;
    double dmax2 = 0.0;
    float lo[3] = {0.0F, 0.0F, 0.0F}, hi[3] = {0.0F, 0.0F, 0.0F};
};
AntennaScan scan_antenna(const float *pos, const float *vel, const float *att, int npri, int ns);

// Host-to-device copies through two page-locked slots on a copy stream of
// their own: put() returns once the source has been read (into a slot) and
// the DMA from that slot is queued; fence() makes a compute stream wait for
// every copy queued so far; begin() makes later copies wait for the work
// queued on a compute stream so far (buffers it still reads).
class Ingest {
  public:
    Ingest() = default;
    Ingest(const Ingest &) = delete;
    Ingest &operator=(const Ingest &) = delete;
    ~Ingest();
    hipError_t begin(hipStream_t compute);
    hipError_t put(void *dst, const void *src, size_t bytes);
    hipError_t fence(hipStream_t compute);
    // later copies wait for `done` (recorded on another stream)
    hipError_t wait(hipEvent_t done);

  private:
    static constexpr size_t kSlotBytes = size_t(4) << 20;
    static constexpr size_t kPieceBytes = size_t(2) << 20;
    static constexpr size_t kParallelCopyBytes = size_t(256) << 10;
    hipError_t init();
    hipStream_t copy_ = nullptr;
    void *slot_[2] = {nullptr, nullptr};
    hipEvent_t free_[2] = {nullptr, nullptr};
    hipEvent_t tail_ = nullptr;
    int cur_ = 0;
    size_t used_ = 0;
    bool ready_ = false;
};

} // namespace tdc
