// tdc_server.cpp -- serving form: continuously batched image formation.
//
// No counterpart in the reference, whose TdcProcessor forms one image per
// blocking start() call (tdc-lib/src/tdcprocessor.cpp:59-167). Two host
// threads drive one GPU:
//   uploaders  (kUploaders threads) take admitted jobs in arrival order, pack
//              each job's inputs into pinned staging memory and move them with
//              one large host-to-device copy (an SDMA transfer, no kernel) on
//              their own stream, run the job's staging kernels behind it, then
//              mark the job resident;
//   scheduler  admits queued jobs in arrival order while their HBM footprint
//              fits the budget, builds steps and completes them.
// A step is a table of tiles (job, pixel tile, chirp range) taken from the
// resident jobs -- an equal share of the step's work for every job first,
// then the rest in arrival order -- plus the finalize blocks of the jobs
// whose last tiles are in it. Steps alternate between two compute streams,
// each replaying the step graph of its grid-size bucket (counter reset + the
// persistent correlate-sum kernel of each tile family it holds), so the next step's
// workgroups fill the CUs while the previous step drains. A job's finalize
// follows the step holding its last tiles on that step's stream, after the
// step before it (on the other stream) has ended too; that covers all of its
// tiles.
//
// Small host-to-device copies would run as blit kernels and wait behind the
// persistent workgroups, so job descriptors and step tables live in pinned
// host memory that the kernels read directly (a few hundred bytes per tile),
// and finished images are written by the finalize straight into pinned host
// memory. The staging kernels of a job (plain records, FP16/BF16 conversion,
// range-profile FFT tables or matrix A tiles) run on its uploader's stream
// right behind its upload, beside the steps of the jobs already resident;
// the job joins the steps once they have run.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <hip/hip_runtime.h>

#include "sarlabtdc_amd.h"
#include "tdc_internal.h"
#include "tdc_serve.h"

namespace {

constexpr double kC = 299792458.0;
constexpr double kDefaultStepTerms = 4e10;
constexpr double kTileTerms = 8388608.0; // ~2^23 work units per tile
// Work units per term of each tile body, one unit being a term of the matrix
// engine (measured on 1024 x 512 x 512^2: matrix 3.70 ms, plain per-chirp range
// 30.6 ms) and, for the range-profile engine, units per (pixel, chirp)
// (0.59 ms there: ~82 matrix terms), so that a step's budget is device time.
constexpr double kPlainTermCost = 8.0;
constexpr double kProfilePairCost = 82.0;
constexpr int kFinWorkgroups = 256;
constexpr int kBuckets[] = {256, 512, 1024, 2048};
constexpr int kNumBuckets = 4;
// Tiles a job is cut into at least (four times the largest persistent grid:
// measured on the 1000-job stream, 2048 gave p99 5.6 / 8.6 / 9.5 ms at 2400 /
// 7200 / 9600 jobs/s and this 2.7 / 6.3 / 7.0 ms), and the shortest chirp
// range of a tile (eight of the range-profile tile's 8-chirp FP32 blocks: on
// the same stream, 64 chirps against 16 cut the saturated makespan by 5% and
// p50 at 9600 jobs/s from 0.87 to 0.76 ms, p50 at 2400 jobs/s unchanged).
constexpr int kFillTiles = 4 * kBuckets[kNumBuckets - 1];
constexpr int kMinSplitChirps = 64;
// The chirp ranges of a job write one partial image each: their bytes are
// kept within this multiple of the job's input bytes (raw samples and grid),
// so that cutting a few-pixel job short does not multiply its footprint.
constexpr double kMaxPartialFactor = 4.0;
// Upload threads: packing a job into pinned memory runs at host memcpy speed,
// so several jobs are packed and copied at once.
constexpr int kUploaders = 8;
// Steps in flight: each has its own tile table, counter and events, so the
// host builds and launches the next steps while earlier ones run and their
// completions wait; steps alternate between two streams.
constexpr int kDepth = 4;

int sfail(int code, const std::string &msg) { return tdc::set_last_error(code, msg); }

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

// Pinned host memory, recycled by power-of-two size class and kept until the
// server is destroyed. Blocks of up to a quarter slab are carved from slabs
// of kSlabBytes, so that a stream of new jobs does not pay one hipHostMalloc
// (25-140 us, measured) per job. Pages are touched when a slab is allocated
// (a job packed into fresh pinned pages was packed at half the speed,
// measured), outside the pool's lock so that other threads keep carving.
class PinnedPool
{
public:
    static constexpr size_t kSlabBytes = size_t(64) << 20;
    explicit PinnedPool(unsigned flags) : flags_(flags) {}
    ~PinnedPool()
    {
        for (void *p : all_) (void) hipHostFree(p);
    }
    // allocates the first slab now rather than on first use
    hipError_t reserve()
    {
        void *p = nullptr;
        const hipError_t e = allocate(kSlabBytes, &p);
        if (e != hipSuccess) return e;
        std::lock_guard<std::mutex> lk(mu_);
        spare_.push_back(p);
        return hipSuccess;
    }
    hipError_t get(size_t n, void **out)
    {
        const size_t c = size_class(n);
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            auto &v = free_[c];
            if (!v.empty()) {
                *out = v.back();
                v.pop_back();
                return hipSuccess;
            }
            if (c > kSlabBytes / 4) {
                lk.unlock();
                return allocate(c, out);
            }
            if (left_ >= c) {
                *out = cur_;
                cur_ += c;
                left_ -= c;
                return hipSuccess;
            }
            if (!spare_.empty()) {
                // the rest of the current slab stays unused
                cur_ = static_cast<char *>(spare_.back());
                spare_.pop_back();
                left_ = kSlabBytes;
                continue;
            }
            lk.unlock();
            void *p = nullptr;
            const hipError_t e = allocate(kSlabBytes, &p);
            lk.lock();
            if (e != hipSuccess) return e;
            spare_.push_back(p);
        }
    }
    void put(void *p, size_t n)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_[size_class(n)].push_back(p);
    }

private:
    static size_t size_class(size_t n)
    {
        size_t c = 4096;
        while (c < n) c <<= 1;
        return c;
    }
    // a new block, its pages touched (without the lock)
    hipError_t allocate(size_t n, void **out)
    {
        const hipError_t e = hipHostMalloc(out, n, flags_);
        if (e != hipSuccess) return e;
        std::memset(*out, 0, n);
        std::lock_guard<std::mutex> lk(mu_);
        all_.push_back(*out);
        return hipSuccess;
    }
    unsigned flags_;
    std::mutex mu_;
    std::map<size_t, std::vector<void *>> free_;
    std::vector<void *> all_;   // every block and slab, freed with the pool
    std::vector<void *> spare_; // slabs not yet carved
    char *cur_ = nullptr;       // the current slab's uncarved part
    size_t left_ = 0;
};

// Device arenas of admitted jobs, by size class (quarter-octave steps from
// 64 KiB, so a class wastes at most 25%), recycled as jobs finish instead of
// being allocated and freed per job: a stream-ordered pool allocation took
// ~0.5 ms per job on the upload threads. Blocks held beyond what the budget
// allows are released before a new one is allocated.
size_t arena_class(size_t n)
{
    size_t base = 65536;
    while (base * 2 <= n) base *= 2;
    if (n <= base) return base;
    for (size_t q = 5; q <= 8; ++q)
        if (base / 4 * q >= n) return base / 4 * q;
    return 2 * base;
}

class DeviceArenas
{
public:
    DeviceArenas() = default;
    DeviceArenas(const DeviceArenas &) = delete;
    DeviceArenas &operator=(const DeviceArenas &) = delete;
    ~DeviceArenas()
    {
        for (auto &kv : size_) (void) hipFree(kv.first);
    }
    // `n` is a class size; `limit` bounds the bytes held (free and in use)
    hipError_t get(size_t n, double limit, char **out)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto &v = free_[n];
            if (!v.empty()) {
                *out = v.back();
                v.pop_back();
                return hipSuccess;
            }
            // release idle blocks of other classes while the new one would not fit
            for (auto it = free_.rbegin(); it != free_.rend() && held_ + n > limit; ++it) {
                while (!it->second.empty() && held_ + n > limit) {
                    char *q = it->second.back();
                    it->second.pop_back();
                    (void) hipFree(q);
                    size_.erase(q);
                    held_ -= it->first;
                }
            }
            held_ += n;
        }
        void *p = nullptr;
        const hipError_t e = hipMalloc(&p, n);
        std::lock_guard<std::mutex> lk(mu_);
        if (e != hipSuccess) {
            held_ -= n;
            return e;
        }
        size_[static_cast<char *>(p)] = n;
        *out = static_cast<char *>(p);
        return hipSuccess;
    }
    // the block's last device use has completed
    void put(char *p, size_t n)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu_);
        free_[n].push_back(p);
    }

private:
    std::mutex mu_;
    std::map<size_t, std::vector<char *>> free_;
    std::unordered_map<char *, size_t> size_;
    size_t held_ = 0;
};

struct Job {
    int64_t id = 0;
    tdc_job_desc d{};
    int state = TDC_JOB_QUEUED;
    int status = TDC_OK;
    std::string error;
    // decomposition (depends only on the job's shape)
    int npx = 0, tile_px = 0, npt = 0, cps = 0, nsplit = 0, nfin = 0;
    int64_t ntiles = 0, next_tile = 0;
    double terms = 0.0;
    double pair_cost = 0.0; // work units per (pixel, chirp)
    // precision jobs: range-profile engine (chirps of at most 8192 samples) or
    // matrix engine
    bool profile = false;
    tdc::ProfileChoice pk; // its kernel and FFT length
    tdc::ProfileParams pp{};
    // device arena layout: [upload region | staged records or raw | partials]
    size_t off_desc = 0, off_cpos = 0, off_ka = 0, off_grid = 0, off_kt = 0, off_ones = 0, off_stage = 0,
           off_pcorr = 0, off_ptw = 0, off_ptab = 0,
           off_atab = 0, off_ascale = 0, off_part = 0, up_bytes = 0, bytes = 0;
    char *arena = nullptr;
    int slot = -1; // admission slot (at most kMaxJobSlots jobs admitted at once)
    const tdc::ServeJob *desc() const
    {
        return reinterpret_cast<const tdc::ServeJob *>(arena + off_desc);
    }
    float *h_image = nullptr; // pinned, written by the finalize kernel, read by wait
    // host
    double t_submit = 0, t_admit = 0, t_first = -1, t_done = 0;
    int steps = 0;
};

struct Step {
    int64_t index = -1;
    int buf = 0;
    std::vector<Job *> finishing;
    std::vector<Job *> touched;
};

} // namespace

struct tdc_server {
    int device = 0;
    double budget = 0.0;
    double step_terms = kDefaultStepTerms;
    hipStream_t step_stream[2] = {nullptr, nullptr}; // steps alternate between these
    hipStream_t up_stream[kUploaders] = {};          // uploads, one per uploader thread
    DeviceArenas arenas; // job arenas, recycled
    hipEvent_t ev_start[kDepth] = {}, ev_end[kDepth] = {};
    hipEvent_t ev_fin[kDepth] = {};
    hipEvent_t ev_origin = nullptr;
    tdc::StepTable *h_table[kDepth] = {};   // pinned step tables
    tdc::StepTable *dev_table[kDepth] = {}; // their device addresses
    unsigned int *d_counter = nullptr; // kDepth pairs of tile counters, 256 bytes apart
    // the value each counter holds when the next step on its table starts
    unsigned int counter_next[kDepth][2] = {};
    // [with the matrix-engine family][grid-size bucket][step table]
    hipGraphExec_t graph[2][kNumBuckets][kDepth] = {};
    int ngraphs = 0;
    // staging for SDMA uploads, and coherent (uncached by the GPU) memory for
    // the images the finalize writes
    PinnedPool pinned{hipHostMallocDefault};
    PinnedPool coherent{hipHostMallocCoherent | hipHostMallocMapped};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_up;
    std::deque<Job *> queued;    // submitted, not yet admitted
    std::deque<Job *> uploading; // admitted, waiting for an uploader
    int uploads_in_progress = 0;
    std::vector<Job *> active;   // resident, tiles outstanding or finishing
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::vector<int> free_slots;
    int64_t next_id = 1;
    bool stop = false;
    bool stop_uploader = false;
    std::string fatal;
    std::thread worker, uploader[kUploaders];

    // stats (guarded by mu)
    int64_t submitted = 0, completed = 0, failed = 0, nsteps = 0;
    double busy_ms = 0.0, terms_done = 0.0, last_end_ms = 0.0;
    size_t in_use = 0, peak = 0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Footprint, arena layout and tile decomposition of a job.
void plan(Job &j)
{
    const tdc_job_desc &d = j.d;
    const size_t nterm = static_cast<size_t>(d.n_pri) * d.n_samples;
    const size_t npx = static_cast<size_t>(d.n_rows) * d.n_cols;
    j.npx = static_cast<int>(npx);
    const bool prec = d.form == TDC_FORM_PRECISION;
    std::string why;
    j.profile = prec && tdc::profile_choice(d.n_samples, static_cast<int64_t>(npx), 0, 0, j.pk, why);
    const int plen = j.pk.M;
    j.pair_cost = j.profile ? kProfilePairCost : (prec ? 1.0 : kPlainTermCost) * d.n_samples;
    j.tile_px = tdc::kServeTilePixels;
    j.npt = (j.npx + j.tile_px - 1) / j.tile_px;
    const double per_chirp = static_cast<double>(std::min(j.npx, j.tile_px)) * j.pair_cost;
    j.cps = std::max(1, static_cast<int>(kTileTerms / std::max(per_chirp, 1.0)));
    // a job with few pixel tiles is cut into shorter chirp ranges until its
    // tiles fill the largest persistent grid (down to kMinSplitChirps): its
    // latency is then that of a short tile, not of one long one. The split
    // depends on the job alone, so its image is the same in any mix.
    const int want_splits = (kFillTiles + j.npt - 1) / j.npt;
    const int fill_cps = std::max(kMinSplitChirps, (d.n_pri + want_splits - 1) / want_splits);
    j.cps = std::min(j.cps, fill_cps);
    j.cps = std::min(j.cps, d.n_pri);
    {
        // FP64 partial images for the matrix engine, FP32 for the others
        const double part_px = (prec && !j.profile) ? sizeof(double2) : sizeof(float2);
        const double input = 8.0 * static_cast<double>(nterm) + (prec ? 24.0 : 12.0) * npx;
        const int max_split =
            std::max(1, static_cast<int>(kMaxPartialFactor * input / (part_px * npx)));
        const int min_cps = (d.n_pri + max_split - 1) / max_split;
        j.cps = std::max(j.cps, std::min(min_cps, d.n_pri));
    }
    j.nsplit = (d.n_pri + j.cps - 1) / j.cps;
    j.ntiles = static_cast<int64_t>(j.npt) * j.nsplit;
    j.nfin = (j.npx + tdc::kBlock - 1) / tdc::kBlock;
    j.terms = static_cast<double>(npx) * d.n_pri * d.n_samples;
    const int nblk = tdc::anchor_entries(d.n_samples);
    size_t o = align256(8 * nterm); // raw samples at offset 0
    j.off_cpos = o;
    o += align256(sizeof(tdc::ChirpPos) * d.n_pri);
    if (prec) {
        j.off_ka = o;
        o += align256(8 * static_cast<size_t>(nblk));
        j.off_grid = o;
        o += align256(24 * npx);
        if (d.raw_storage != TDC_RAW_C64) {
            j.off_ones = o;
            o += align256(4 * static_cast<size_t>(d.n_samples));
        }
        if (j.profile) {
            j.off_pcorr = o;
            o += align256(4 * static_cast<size_t>(d.n_samples));
            j.off_ptw = o;
            o += align256(sizeof(float2) * static_cast<size_t>(plen / 2));
        }
    } else {
        j.off_kt = o;
        o += align256(4 * static_cast<size_t>(d.n_samples));
        j.off_ka = o;
        o += align256(8 * static_cast<size_t>((d.n_samples + tdc::kAnchorBlock - 1) /
                                              tdc::kAnchorBlock));
        j.off_grid = o;
        o += align256(24 * npx);
    }
    j.off_desc = o; // the job's descriptor (device copy read by the kernels)
    o += align256(sizeof(tdc::ServeJob));
    j.up_bytes = o;
    j.off_stage = o;
    if (!prec) {
        o += align256(sizeof(float4) * nterm);
    } else if (d.raw_storage != TDC_RAW_C64) {
        o += align256(4 * nterm);
    }
    if (j.profile) {
        j.off_ptab = o;
        o += align256(sizeof(float2) * static_cast<size_t>(d.n_pri) * (plen + tdc::kProfilePad));
    } else if (prec) {
        j.off_atab = o;
        o += align256(tdc::matrix_table_bytes(d.n_pri, d.n_samples, d.raw_storage));
        j.off_ascale = o;
        o += align256(sizeof(float) * d.n_pri);
    }
    j.off_part = o;
    o += align256(((prec && !j.profile) ? sizeof(double2) : sizeof(float2)) * npx * j.nsplit);
    j.bytes = arena_class(o);
}

int validate(const tdc_job_desc *d)
{
    if (!d) return sfail(TDC_ERR_INVALID_ARGUMENT, "job description is NULL");
    if (d->n_pri <= 0) return sfail(TDC_ERR_INVALID_ARGUMENT, "nPri must be greater than 0");
    if (d->n_samples <= 0) return sfail(TDC_ERR_INVALID_ARGUMENT, "nSamples must be greater than 0");
    if (d->n_rows <= 0 || d->n_cols <= 0)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "the focus grid must not be empty");
    if (static_cast<int64_t>(d->n_pri) * d->n_samples > INT_MAX)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "nPri * nSamples must not exceed 2^31 - 1");
    if (static_cast<int64_t>(d->n_rows) * d->n_cols > tdc::kMaxPixels)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "the focus grid must not exceed 2^30 pixels");
    if (!d->raw_iq || !d->sample_times || !d->chirp_position || !d->grid)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "job arrays must not be NULL");
    if (d->form != TDC_FORM_PLAIN && d->form != TDC_FORM_PRECISION)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "unknown form");
    if (d->raw_storage < TDC_RAW_C64 || d->raw_storage > TDC_RAW_BF16)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "unknown raw storage");
    if (d->form == TDC_FORM_PLAIN && d->raw_storage != TDC_RAW_C64)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "the plain form reads complex64 raw data");
    return TDC_OK;
}

#define SRV_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            err = std::string(#expr) + " failed: " + hipGetErrorString(e_);                   \
            return false;                                                                      \
        }                                                                                      \
    } while (0)

hipError_t stage_job(const Job &j, hipStream_t ss);

// Copies npx FP64 points [npx][3] and returns their bounding box (NaN bounds
// when a coordinate is NaN), in one pass over the source.
void copy_grid_bounds(double *dst, const double *src, size_t npx, double lo[3], double hi[3])
{
    const double inf = std::numeric_limits<double>::infinity();
    double l0 = inf, l1 = inf, l2 = inf, h0 = -inf, h1 = -inf, h2 = -inf;
    bool nan = false;
    for (size_t i = 0; i < npx; ++i) {
        const double x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2];
        dst[3 * i] = x;
        dst[3 * i + 1] = y;
        dst[3 * i + 2] = z;
        nan = nan || x != x || y != y || z != z;
        l0 = std::min(l0, x);
        l1 = std::min(l1, y);
        l2 = std::min(l2, z);
        h0 = std::max(h0, x);
        h1 = std::max(h1, y);
        h2 = std::max(h2, z);
    }
    const double q = std::numeric_limits<double>::quiet_NaN();
    lo[0] = nan ? q : l0;
    lo[1] = nan ? q : l1;
    lo[2] = nan ? q : l2;
    hi[0] = nan ? q : h0;
    hi[1] = nan ? q : h1;
    hi[2] = nan ? q : h2;
}

// Pack, upload, stage and describe one admitted job (uploader thread).
bool upload_job(tdc_server *s, Job &j, hipStream_t us, std::string &err)
{
    tdc::Range range("tdc_server_upload");
    const tdc_job_desc &d = j.d;
    const size_t nterm = static_cast<size_t>(d.n_pri) * d.n_samples;
    const bool prec = d.form == TDC_FORM_PRECISION;
    std::vector<double> kanchor;
    double dk = 0.0;
    if (prec && !tdc::precision_tables(d.sample_times, d.n_samples, d.start_freq, d.mod_rate,
                                       kanchor, dk, err))
        return false;
    void *stage_mem = nullptr;
    SRV_HIP(s->pinned.get(j.up_bytes, &stage_mem));
    char *h = static_cast<char *>(stage_mem);
    std::memcpy(h, d.raw_iq, 8 * nterm);
    auto *hc = reinterpret_cast<tdc::ChirpPos *>(h + j.off_cpos);
    for (int n = 0; n < d.n_pri; ++n) {
        const double *c = d.chirp_position + 3 * static_cast<size_t>(n);
        hc[n] = tdc::ChirpPos{c[0], c[1], c[2], 0.0};
    }
    // the grid's bounding box comes with its copy (range-profile jobs)
    double glo[3], ghi[3];
    if (prec) {
        std::memcpy(h + j.off_ka, kanchor.data(), 8 * kanchor.size());
        if (j.profile)
            copy_grid_bounds(reinterpret_cast<double *>(h + j.off_grid), d.grid,
                             static_cast<size_t>(j.npx), glo, ghi);
        else
            std::memcpy(h + j.off_grid, d.grid, 24 * static_cast<size_t>(j.npx));
        if (d.raw_storage != TDC_RAW_C64) {
            // served raw data arrive windowed: the conversion uses a unit window
            auto *w = reinterpret_cast<float *>(h + j.off_ones);
            for (int m = 0; m < d.n_samples; ++m) w[m] = 1.0F;
        }
        if (j.profile) {
            // the processor's range-profile set-up: chirp positions from an
            // origin inside the scene with their squared norms, the kernel's
            // correction and the FFT's twiddles
            const int ns = d.n_samples;
            const int plen = j.pk.M;
            const size_t last = 3 * (static_cast<size_t>(j.npx) - 1);
            double o[3];
            for (int a = 0; a < 3; ++a) {
                o[a] = 0.5 * (d.grid[a] + d.grid[last + a]);
                if (!std::isfinite(o[a])) o[a] = 0.0;
            }
            for (int n = 0; n < d.n_pri; ++n) {
                const double x = hc[n].x - o[0], y = hc[n].y - o[1], z = hc[n].z - o[2];
                hc[n] = tdc::ChirpPos{x, y, z, std::fma(z, z, std::fma(y, y, x * x))};
            }
            const auto tabs = tdc::profile_tables(ns, plen, j.pk.taps, j.pk.over);
            std::memcpy(h + j.off_pcorr, tabs->corr.data(), 4 * tabs->corr.size());
            std::memcpy(h + j.off_ptw, tabs->twiddle.data(), sizeof(float2) * tabs->twiddle.size());
            const double dt = ns > 1 ? (d.sample_times[ns - 1] - d.sample_times[0]) / (ns - 1) : 0.0;
            j.pp.dk = dk;
            j.pp.kh = 2.0 * (d.start_freq + d.mod_rate * (d.sample_times[0] + (ns / 2) * dt)) / kC;
            j.pp.ox = o[0];
            j.pp.oy = o[1];
            j.pp.oz = o[2];
            j.pp.M = plen;
            j.pp.Mp = plen + tdc::kProfilePad;
            j.pp.taps = j.pk.taps;
            j.pp.over = j.pk.over;
            j.pp.dkm = std::fabs(dk) * plen;
            j.pp.reverse = dk < 0.0 ? 1 : 0;
            for (int a = 0; a < 3; ++a) {
                glo[a] -= o[a];
                ghi[a] -= o[a];
            }
            j.pp.incr = tdc::profile_incremental_ok(hc, d.n_pri, glo, ghi,
                                                    std::max(std::fabs(j.pp.kh), j.pp.dkm))
                            ? 1
                            : 0;
        }
    } else {
        // the plain form's phase tables (the processor's, from FP64 inputs) and
        // the FP64 grid: the range is computed in FP64 from FP64 coordinates
        std::vector<double> k(d.n_samples), kb;
        std::vector<float> dkm;
        for (int m = 0; m < d.n_samples; ++m)
            k[m] = 2.0 * (d.start_freq + d.mod_rate * d.sample_times[m]) / kC;
        tdc::plain_phase_tables(k, dkm, kb);
        std::memcpy(h + j.off_kt, dkm.data(), 4 * dkm.size());
        std::memcpy(h + j.off_ka, kb.data(), 8 * kb.size());
        std::memcpy(h + j.off_grid, d.grid, 24 * static_cast<size_t>(j.npx));
    }
    char *arena = nullptr;
    hipError_t e = s->arenas.get(j.bytes, s->budget, &arena);
    void *img = nullptr, *img_dev = nullptr;
    if (e == hipSuccess) {
        j.arena = arena;
        e = s->coherent.get(sizeof(float2) * j.npx, &img);
    }
    if (e == hipSuccess) {
        j.h_image = static_cast<float *>(img);
        e = hipHostGetDevicePointer(&img_dev, img, 0);
    }
    if (e != hipSuccess) {
        s->pinned.put(stage_mem, j.up_bytes);
        err = std::string("job set-up failed: ") + hipGetErrorString(e);
        return false;
    }
    // the descriptor the kernels read, with the job's device addresses; it
    // travels in the job's upload and tiles point at its device copy
    tdc::ServeJob sj{};
    sj.cpos = reinterpret_cast<const tdc::ChirpPos *>(j.arena + j.off_cpos);
    sj.image = static_cast<float2 *>(img_dev);
    sj.npx = j.npx;
    sj.npri = d.n_pri;
    sj.ns = d.n_samples;
    sj.nsplit = j.nsplit;
    sj.cps = j.cps;
    sj.npt = j.npt;
    sj.form = prec ? 1 : 0;
    sj.storage = d.raw_storage;
    sj.partial = j.arena + j.off_part;
    if (j.profile) {
        sj.raw = j.arena + j.off_ptab;
        sj.grid64 = reinterpret_cast<const double *>(j.arena + j.off_grid);
        sj.pp = j.pp;
        sj.profile = 1;
    } else if (prec) {
        sj.raw = j.arena + j.off_atab;
        sj.inv_scale = reinterpret_cast<const float *>(j.arena + j.off_ascale);
        sj.grid64 = reinterpret_cast<const double *>(j.arena + j.off_grid);
        sj.kanchor = reinterpret_cast<const double *>(j.arena + j.off_ka);
        sj.dk = dk;
    } else {
        sj.raw = j.arena + j.off_stage;
        sj.grid64 = reinterpret_cast<const double *>(j.arena + j.off_grid);
        sj.kanchor = reinterpret_cast<const double *>(j.arena + j.off_ka);
    }
    std::memcpy(h + j.off_desc, &sj, sizeof(sj));
    e = hipMemcpyAsync(arena, stage_mem, j.up_bytes, hipMemcpyHostToDevice, us);
    // the job's staging kernels (FFT tables, A tiles, records) right behind its
    // upload on this thread's stream: they overlap the steps of other jobs,
    // and the job is resident once they have run
    if (e == hipSuccess) e = stage_job(j, us);
    if (e == hipSuccess) e = hipStreamSynchronize(us);
    s->pinned.put(stage_mem, j.up_bytes);
    if (e != hipSuccess) {
        err = std::string("job upload failed: ") + hipGetErrorString(e);
        return false;
    }
    return true;
}

// Staging kernels of a job, on its uploader's stream right behind its upload.
hipError_t stage_job(const Job &j, hipStream_t ss)
{
    const tdc_job_desc &d = j.d;
    if (d.form == TDC_FORM_PRECISION) {
        // raw samples in their storage format (complex64 as uploaded), then
        // the range-profile table or the matrix engine's A tiles
        const void *staged = j.arena;
        if (d.raw_storage != TDC_RAW_C64) {
            hipError_t e = tdc::launch_stage_precision(
                d.raw_storage, reinterpret_cast<const float *>(j.arena),
                reinterpret_cast<const float *>(j.arena + j.off_ones), d.n_pri, d.n_samples,
                j.arena + j.off_stage, ss);
            if (e != hipSuccess) return e;
            staged = j.arena + j.off_stage;
        }
        if (j.profile)
            return tdc::launch_stage_profile(
                d.raw_storage, staged, reinterpret_cast<const float *>(j.arena + j.off_pcorr),
                reinterpret_cast<const float2 *>(j.arena + j.off_ptw),
                reinterpret_cast<const tdc::ChirpPos *>(j.arena + j.off_cpos), d.n_pri,
                d.n_samples, j.pp.M, j.pp.Mp, j.pp.taps, j.pp.reverse,
                reinterpret_cast<float2 *>(j.arena + j.off_ptab), ss);
        return tdc::launch_stage_matrix(d.raw_storage, staged, d.n_pri, d.n_samples,
                                        tdc::matrix_parts(d.raw_storage), j.arena + j.off_atab,
                                        reinterpret_cast<float *>(j.arena + j.off_ascale), ss);
    }
    return tdc::launch_stage_plain_chirp(reinterpret_cast<const float *>(j.arena),
                                         reinterpret_cast<const float *>(j.arena + j.off_kt),
                                         d.n_pri, d.n_samples,
                                         reinterpret_cast<float4 *>(j.arena + j.off_stage), ss);
}

bool ensure_graph(tdc_server *s, int matrix, int bucket, int buf, std::string &err)
{
    if (s->graph[matrix][bucket][buf]) return true;
    hipStream_t cs = nullptr;
    SRV_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        e = tdc::launch_serve_step(s->dev_table[buf], s->d_counter + 64 * buf, kBuckets[bucket],
                                   matrix != 0, cs);
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&s->graph[matrix][bucket][buf], g, nullptr, nullptr, 0);
    if (g) (void) hipGraphDestroy(g);
    (void) hipStreamDestroy(cs);
    if (e != hipSuccess) {
        err = std::string("step graph capture failed: ") + hipGetErrorString(e);
        return false;
    }
    ++s->ngraphs;
    return true;
}

void fail_job_locked(tdc_server *s, Job *j, int code, const std::string &msg)
{
    j->state = TDC_JOB_FAILED;
    j->status = code;
    j->error = msg;
    j->t_done = s->now();
    ++s->failed;
}

// Finish the jobs of a step whose finalize has ended (scheduler thread).
bool complete_step(tdc_server *s, Step &st, std::string &err)
{
    tdc::Range range("tdc_server_complete_step");
    SRV_HIP(hipEventSynchronize(s->ev_fin[st.buf]));
    float t0 = 0.0F, t1 = 0.0F;
    SRV_HIP(hipEventElapsedTime(&t0, s->ev_origin, s->ev_start[st.buf]));
    SRV_HIP(hipEventElapsedTime(&t1, s->ev_origin, s->ev_end[st.buf]));
    for (Job *j : st.finishing) {
        s->arenas.put(j->arena, j->bytes);
        j->arena = nullptr;
    }
    const double t = s->now();
    std::lock_guard<std::mutex> lk(s->mu);
    // device busy time: union of the step intervals (consecutive steps overlap)
    const double a = std::max(static_cast<double>(t0), s->last_end_ms);
    if (t1 > a) s->busy_ms += t1 - a;
    s->last_end_ms = std::max(s->last_end_ms, static_cast<double>(t1));
    for (Job *j : st.finishing) {
        j->state = TDC_JOB_DONE;
        j->t_done = t;
        s->in_use -= j->bytes;
        s->free_slots.push_back(j->slot);
        s->terms_done += j->terms;
        ++s->completed;
        s->active.erase(std::find(s->active.begin(), s->active.end(), j));
    }
    s->cv_done.notify_all();
    s->cv_sched.notify_all();
    return true;
}

// Uploads admitted jobs, taken in arrival order (uploader thread `k`); a job
// becomes resident when its own upload has ended.
void uploader(tdc_server *s, int k)
{
    if (hipSetDevice(s->device) != hipSuccess) return;
    for (;;) {
        Job *j = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_up.wait(lk, [&] { return !s->uploading.empty() || s->stop_uploader; });
            if (s->uploading.empty()) return;
            j = s->uploading.front();
            s->uploading.pop_front();
            ++s->uploads_in_progress;
        }
        std::string why;
        const bool okay = upload_job(s, *j, s->up_stream[k], why);
        std::lock_guard<std::mutex> lk(s->mu);
        --s->uploads_in_progress;
        if (okay) {
            s->active.push_back(j);
        } else {
            s->in_use -= j->bytes;
            s->free_slots.push_back(j->slot);
            if (j->arena) {
                (void) hipStreamSynchronize(s->up_stream[k]);
                s->arenas.put(j->arena, j->bytes);
            }
            j->arena = nullptr;
            if (j->h_image) s->coherent.put(j->h_image, sizeof(float2) * j->npx);
            j->h_image = nullptr;
            fail_job_locked(s, j, TDC_ERR_UNSUPPORTED, why);
            s->cv_done.notify_all();
        }
        s->cv_sched.notify_all();
    }
}

void scheduler(tdc_server *s)
{
    std::string err;
    std::deque<Step> inflight;
    int64_t step_index = 0;
    auto die = [&](const std::string &msg) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->fatal = msg;
        for (auto &kv : s->jobs) {
            Job *j = kv.second.get();
            if (j->state == TDC_JOB_QUEUED || j->state == TDC_JOB_ADMITTED)
                fail_job_locked(s, j, TDC_ERR_DEVICE, msg);
        }
        s->queued.clear();
        s->active.clear();
        s->cv_done.notify_all();
    };
    if (hipSetDevice(s->device) != hipSuccess) {
        die("hipSetDevice failed in the scheduler thread");
        return;
    }
    for (;;) {
        // 1. admission in arrival order: charge the budget, hand to the uploader
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->queued.empty() && s->uploading.empty() && s->uploads_in_progress == 0 &&
                s->active.empty() && inflight.empty()) {
                if (s->stop) break;
                s->cv_sched.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
            while (!s->queued.empty()) {
                Job *j = s->queued.front();
                if (s->in_use + j->bytes > s->budget || s->free_slots.empty()) break;
                s->queued.pop_front();
                s->in_use += j->bytes;
                s->peak = std::max(s->peak, s->in_use);
                j->slot = s->free_slots.back();
                s->free_slots.pop_back();
                j->state = TDC_JOB_ADMITTED;
                j->t_admit = s->now();
                s->uploading.push_back(j);
                s->cv_up.notify_one();
            }
        }

        // 2. completions: the oldest step when two are in flight (its table
        //    buffer is needed next), or any step whose finalize already ended
        while (!inflight.empty() &&
               (inflight.size() == static_cast<size_t>(kDepth) ||
                hipEventQuery(s->ev_fin[inflight.front().buf]) == hipSuccess)) {
            if (!complete_step(s, inflight.front(), err)) return die(err);
            inflight.pop_front();
        }

        // 3. build the next step from the resident jobs
        Step st;
        st.index = step_index;
        st.buf = static_cast<int>(step_index % kDepth);
        tdc::StepTable *tab = s->h_table[st.buf];
        int nt = 0, nf = 0;
        // runs of consecutive tiles of one job, per tile family: 0 plain and
        // range-profile tiles, 1 matrix-engine tiles (their own persistent kernel)
        struct Run {
            const tdc::ServeJob *job;
            int first, count;
            int64_t tile0;
        };
        std::vector<Run> runs[2];
        int nfam[2] = {0, 0};
        double work = 0.0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            // processor sharing: every resident job with tiles left gets an equal
            // share of the step (at least one tile), oldest first; what the
            // shares leave is then filled in arrival order
            int nwant = 0;
            for (Job *j : s->active)
                if (j->next_tile < j->ntiles) ++nwant;
            const double share = nwant > 0 ? s->step_terms / nwant : 0.0;
            auto take_tiles = [&](Job *j, double limit) {
                double got = 0.0;
                bool took = false;
                const int fam = (j->d.form == TDC_FORM_PRECISION && !j->profile) ? 1 : 0;
                std::vector<Run> &rr = runs[fam];
                // the job's tiles continue its last run of this step, or open a new
                // run while the step has room for one
                const bool extend = !rr.empty() && rr.back().job == j->desc() &&
                                    rr.back().tile0 + rr.back().count == j->next_tile;
                if (!extend && runs[0].size() + runs[1].size() >= size_t(tdc::kMaxStepRuns))
                    return;
                if (!extend) rr.push_back(Run{j->desc(), nfam[fam], 0, j->next_tile});
                while (j->next_tile < j->ntiles && nt < tdc::kMaxStepTiles &&
                       (got < limit || !took) && (work < s->step_terms || nt == 0)) {
                    const int64_t t = j->next_tile++;
                    const int split = static_cast<int>(t / j->npt);
                    const int ptile = static_cast<int>(t % j->npt);
                    ++rr.back().count;
                    ++nfam[fam];
                    ++nt;
                    const int n0 = split * j->cps;
                    const int nch = std::min(j->d.n_pri, n0 + j->cps) - n0;
                    const int npx = std::min(j->tile_px, j->npx - ptile * j->tile_px);
                    const double w = static_cast<double>(npx) * nch * j->pair_cost;
                    work += w;
                    got += w;
                    took = true;
                }
                if (!took && !extend) rr.pop_back();
                if (took) {
                    if (j->t_first < 0) {
                        j->t_first = s->now();
                    }
                    if (std::find(st.touched.begin(), st.touched.end(), j) == st.touched.end()) {
                        ++j->steps;
                        st.touched.push_back(j);
                    }
                }
            };
            for (Job *j : s->active) {
                if (work >= s->step_terms || nt >= tdc::kMaxStepTiles) break;
                take_tiles(j, share);
            }
            for (Job *j : s->active) {
                if (work >= s->step_terms || nt >= tdc::kMaxStepTiles) break;
                take_tiles(j, s->step_terms);
            }
            for (Job *j : s->active) {
                // jobs whose last tile is in this step (or an earlier one) finish here
                if (j->next_tile == j->ntiles && nf + j->nfin <= tdc::kMaxStepFin) {
                    for (int f = 0; f < j->nfin; ++f)
                        tab->fin[nf++] = tdc::FinTile{j->desc(), f * tdc::kBlock, 0};
                    st.finishing.push_back(j);
                    j->next_tile = j->ntiles + 1; // finalize scheduled
                }
            }
        }
        if (nt == 0 && nf == 0) {
            if (!inflight.empty()) {
                if (!complete_step(s, inflight.front(), err)) return die(err);
                inflight.pop_front();
            } else {
                std::unique_lock<std::mutex> lk(s->mu);
                s->cv_sched.wait_for(lk, std::chrono::milliseconds(2));
            }
            continue;
        }
        {
            int r = 0;
            for (int fam = 0; fam < 2; ++fam) {
                tab->ntiles[fam] = nfam[fam];
                tab->nruns[fam] = static_cast<int>(runs[fam].size());
                for (const Run &run : runs[fam]) {
                    tab->runs[r] = tdc::StepRun{run.job, run.first, run.count};
                    tab->tile0[r] = static_cast<int>(run.tile0);
                    ++r;
                }
            }
        }
        tab->nfin = nf;
        tab->base[0] = s->counter_next[st.buf][0];
        tab->base[1] = s->counter_next[st.buf][1];
        int bucket = kNumBuckets - 1;
        for (int b = 0; b < kNumBuckets; ++b) {
            if (kBuckets[b] >= nt) {
                bucket = b;
                break;
            }
        }
        // a step without matrix-engine tiles replays the graph without their
        // (otherwise empty) persistent kernel
        const int matrix = nfam[1] > 0 ? 1 : 0;
        if (!ensure_graph(s, matrix, bucket, st.buf, err)) return die(err);
        tdc::Range range("tdc_server_launch_step");
        const int p = st.buf;
        const int prev = (p + kDepth - 1) % kDepth; // the step before, on the other stream
        hipStream_t ss = s->step_stream[p & 1];
        hipError_t e = hipEventRecord(s->ev_start[p], ss);
        if (e == hipSuccess) e = hipGraphLaunch(s->graph[matrix][bucket][p], ss);
        // every workgroup of a family's kernel takes one counter value past the
        // family's tiles
        s->counter_next[p][0] += static_cast<unsigned int>(nfam[0] + kBuckets[bucket]);
        if (matrix) s->counter_next[p][1] += static_cast<unsigned int>(nfam[1] + kBuckets[bucket]);
        if (e == hipSuccess) e = hipEventRecord(s->ev_end[p], ss);
        // finalize on the same stream, right behind this step, once the previous
        // step (the other stream) has ended too: the two cover every earlier
        // step of the finishing jobs
        if (e == hipSuccess && step_index > 0 && nf > 0)
            e = hipStreamWaitEvent(ss, s->ev_end[prev], 0);
        if (e == hipSuccess && nf > 0)
            e = tdc::launch_serve_finalize(s->dev_table[p], kFinWorkgroups, ss);
        if (e == hipSuccess) e = hipEventRecord(s->ev_fin[p], ss);
        if (e != hipSuccess) return die(std::string("step launch failed: ") + hipGetErrorString(e));
        {
            std::lock_guard<std::mutex> lk(s->mu);
            ++s->nsteps;
        }
        ++step_index;
        inflight.push_back(std::move(st));
    }
    while (!inflight.empty()) {
        if (!complete_step(s, inflight.front(), err)) return die(err);
        inflight.pop_front();
    }
}

} // namespace

extern "C" {

int tdc_server_create(int gpu_num, double hbm_budget_bytes, double step_terms, tdc_server **out)
{
    if (!out) return sfail(TDC_ERR_INVALID_ARGUMENT, "out handle pointer is NULL");
    *out = nullptr;
    if (!(hbm_budget_bytes > 0.0)) return sfail(TDC_ERR_INVALID_ARGUMENT, "the HBM budget must be positive");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || gpu_num < 0 || gpu_num >= n)
        return sfail(TDC_ERR_DEVICE, "Failed to set HIP device " + std::to_string(gpu_num));
    auto s = std::make_unique<tdc_server>();
    s->device = gpu_num;
    s->budget = hbm_budget_bytes;
    if (step_terms > 0.0) s->step_terms = step_terms;
    std::string err;
    auto init = [&]() -> bool {
        SRV_HIP(hipSetDevice(gpu_num));
        for (auto &us : s->up_stream) SRV_HIP(hipStreamCreateWithFlags(&us, hipStreamNonBlocking));
        SRV_HIP(s->pinned.reserve());
        SRV_HIP(s->coherent.reserve());
        SRV_HIP(hipEventCreate(&s->ev_origin));
        for (auto &ss : s->step_stream) SRV_HIP(hipStreamCreateWithFlags(&ss, hipStreamNonBlocking));
        for (int i = 0; i < kDepth; ++i) {
            SRV_HIP(hipEventCreate(&s->ev_start[i]));
            SRV_HIP(hipEventCreate(&s->ev_end[i]));
            SRV_HIP(hipEventCreateWithFlags(&s->ev_fin[i], hipEventDisableTiming));
            // step tables and the job table are rewritten by the host between
            // launches: coherent memory, so no GPU cache keeps a stale copy
            SRV_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_table[i]), sizeof(tdc::StepTable),
                                  hipHostMallocCoherent | hipHostMallocMapped));
            SRV_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&s->dev_table[i]),
                                            s->h_table[i], 0));
        }
        SRV_HIP(hipMalloc(reinterpret_cast<void **>(&s->d_counter), 256 * kDepth));
        // the tile counters start at zero and are never reset (steps read them
        // relative to the bases in their tables)
        SRV_HIP(hipMemset(s->d_counter, 0, 256 * kDepth));
        SRV_HIP(hipDeviceSynchronize());
        SRV_HIP(hipEventRecord(s->ev_origin, s->step_stream[0]));
        SRV_HIP(hipEventSynchronize(s->ev_origin));
        return true;
    };
    if (!init()) {
        tdc_server_destroy(s.release());
        return sfail(TDC_ERR_DEVICE, err);
    }
    for (int i = tdc::kMaxJobSlots - 1; i >= 0; --i) s->free_slots.push_back(i);
    tdc_server *raw = s.release();
    raw->worker = std::thread(scheduler, raw);
    for (int k = 0; k < kUploaders; ++k) raw->uploader[k] = std::thread(uploader, raw, k);
    *out = raw;
    tdc::clear_last_error();
    return TDC_OK;
}

int tdc_server_submit(tdc_server *s, const tdc_job_desc *job, int64_t *job_id)
{
    if (!s || !job_id) return sfail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    int rc = validate(job);
    if (rc != TDC_OK) return rc;
    auto j = std::make_unique<Job>();
    j->d = *job;
    plan(*j);
    if (static_cast<double>(j->bytes) > s->budget)
        return sfail(TDC_ERR_OUT_OF_MEMORY, "the job needs " + std::to_string(j->bytes) +
                                                " bytes of HBM, more than the whole budget");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->fatal.empty()) return sfail(TDC_ERR_DEVICE, s->fatal);
    if (s->stop) return sfail(TDC_ERR_STATE, "the server is shutting down");
    j->id = s->next_id++;
    j->t_submit = s->now();
    *job_id = j->id;
    s->queued.push_back(j.get());
    s->jobs[j->id] = std::move(j);
    ++s->submitted;
    s->cv_sched.notify_one();
    tdc::clear_last_error();
    return TDC_OK;
}

int tdc_server_poll(tdc_server *s, int64_t job_id, int *done)
{
    if (!s || !done) return sfail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return sfail(TDC_ERR_INVALID_ARGUMENT, "unknown job id");
    const int st = it->second->state;
    *done = (st == TDC_JOB_DONE || st == TDC_JOB_FAILED) ? 1 : 0;
    return TDC_OK;
}

int tdc_server_wait(tdc_server *s, int64_t job_id, float *image_iq, tdc_job_info *info)
{
    if (!s) return sfail(TDC_ERR_INVALID_ARGUMENT, "NULL server");
    std::unique_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job_id);
        if (it == s->jobs.end()) return sfail(TDC_ERR_INVALID_ARGUMENT, "unknown job id");
        Job *p = it->second.get();
        s->cv_done.wait(lk, [&] { return p->state == TDC_JOB_DONE || p->state == TDC_JOB_FAILED; });
        j = std::move(it->second);
        s->jobs.erase(it);
    }
    if (info) {
        info->id = j->id;
        info->state = j->state;
        info->status = j->status;
        info->t_submit_s = j->t_submit;
        info->t_admit_s = j->t_admit;
        info->t_first_step_s = j->t_first;
        info->t_done_s = j->t_done;
        info->terms = j->terms;
        info->device_bytes = static_cast<int64_t>(j->bytes);
        info->tiles = static_cast<int32_t>(j->ntiles);
        info->steps = j->steps;
    }
    if (j->state == TDC_JOB_FAILED) return sfail(j->status, j->error);
    if (image_iq) std::memcpy(image_iq, j->h_image, sizeof(float2) * static_cast<size_t>(j->npx));
    s->coherent.put(j->h_image, sizeof(float2) * j->npx);
    j->h_image = nullptr;
    tdc::clear_last_error();
    return TDC_OK;
}

int tdc_server_get_stats(tdc_server *s, tdc_server_stats *out)
{
    if (!s || !out) return sfail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->steps = s->nsteps;
    out->device_busy_ms = s->busy_ms;
    out->terms_done = s->terms_done;
    out->bytes_in_use = static_cast<int64_t>(s->in_use);
    out->bytes_peak = static_cast<int64_t>(s->peak);
    out->queued = static_cast<int32_t>(s->queued.size());
    out->active = static_cast<int32_t>(s->active.size());
    out->graphs = s->ngraphs;
    return TDC_OK;
}

int tdc_server_clock(tdc_server *s, double *seconds)
{
    if (!s || !seconds) return sfail(TDC_ERR_INVALID_ARGUMENT, "NULL argument");
    *seconds = s->now();
    return TDC_OK;
}

int tdc_server_destroy(tdc_server *s)
{
    if (!s) return TDC_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Job *j : s->queued) fail_job_locked(s, j, TDC_ERR_STATE, "the server shut down");
        s->queued.clear();
        s->cv_sched.notify_all();
        s->cv_done.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop_uploader = true;
        s->cv_up.notify_all();
    }
    for (auto &t : s->uploader)
        if (t.joinable()) t.join();
    (void) hipSetDevice(s->device);
    (void) hipDeviceSynchronize();
    for (auto &fam : s->graph)
        for (auto &row : fam)
            for (auto &g : row)
                if (g) (void) hipGraphExecDestroy(g);
    for (int i = 0; i < kDepth; ++i) {
        if (s->ev_start[i]) (void) hipEventDestroy(s->ev_start[i]);
        if (s->ev_end[i]) (void) hipEventDestroy(s->ev_end[i]);
        if (s->ev_fin[i]) (void) hipEventDestroy(s->ev_fin[i]);
        if (s->h_table[i]) (void) hipHostFree(s->h_table[i]);
    }
    for (auto &ss : s->step_stream)
        if (ss) (void) hipStreamDestroy(ss);
    if (s->ev_origin) (void) hipEventDestroy(s->ev_origin);
    if (s->d_counter) (void) hipFree(s->d_counter);
    for (auto &us : s->up_stream)
        if (us) (void) hipStreamDestroy(us);
    delete s;
    return TDC_OK;
}

} // extern "C"
