// tdc_server.cpp -- serving form: continuously batched image formation.
//
// No counterpart in the reference, whose TdcProcessor forms one image per
// blocking start() call (tdc-lib/src/tdcprocessor.cpp:59-167). Here a
// scheduler thread owns one HIP stream and:
//   1. admits queued jobs in arrival order while their HBM footprint fits the
//      budget, uploads their inputs and stages them (stream-ordered
//      allocations, so nothing waits for the device);
//   2. builds a step: tiles (job, pixel tile, chirp range) taken from the
//      admitted jobs in arrival order up to a work budget, plus the finalize
//      blocks of the jobs whose last tiles are in the step;
//   3. copies the step table to the device and replays the step graph of the
//      grid-size bucket (reset counter, persistent correlate-sum kernel,
//      finalize), keeping two steps in flight;
//   4. when a step's event completes, copies the finished images to host
//      memory, frees their device memory and wakes their waiters.
// New jobs join at the next step boundary while earlier ones are still
// running, and their tiles share the device with them.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_runtime.h>

#include "sarlabtdc_amd.h"
#include "tdc_internal.h"
#include "tdc_serve.h"

namespace {

constexpr double kC = 299792458.0;
constexpr double kDefaultStepTerms = 4e10;
constexpr double kTileTerms = 8388608.0; // ~2^23 terms per tile
constexpr int kFinWorkgroups = 256;
constexpr int kBuckets[] = {256, 512, 1024, 2048};
constexpr int kNumBuckets = 4;

int sfail(int code, const std::string &msg) { return tdc::set_last_error(code, msg); }

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

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
    size_t bytes = 0;
    // device
    void *arena = nullptr;
    int slot = -1;
    float2 *d_image = nullptr;
    // host
    std::vector<float> image;
    double t_submit = 0, t_admit = 0, t_first = -1, t_done = 0;
    int steps = 0;
    int64_t last_step = -1;
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
    hipStream_t stream = nullptr, copy_stream = nullptr;
    hipEvent_t ev_start[2] = {nullptr, nullptr}, ev_end[2] = {nullptr, nullptr};
    hipEvent_t ev_upload = nullptr;
    tdc::ServeJob *d_jobs = nullptr;
    tdc::StepTable *d_table[2] = {nullptr, nullptr};
    tdc::StepTable *h_table[2] = {nullptr, nullptr};
    unsigned int *d_counter = nullptr;
    hipGraphExec_t graph[kNumBuckets][2] = {};
    int ngraphs = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<Job *> queued;
    std::vector<Job *> active;            // admitted, tiles outstanding or finishing
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::vector<int> free_slots;
    int64_t next_id = 1;
    bool stop = false;
    std::string fatal;
    std::thread worker;

    // stats (guarded by mu)
    int64_t submitted = 0, completed = 0, failed = 0, nsteps = 0;
    double busy_ms = 0.0, terms_done = 0.0;
    size_t in_use = 0, peak = 0;

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Footprint and tile decomposition of a job.
void plan(Job &j)
{
    const tdc_job_desc &d = j.d;
    const size_t nterm = static_cast<size_t>(d.n_pri) * d.n_samples;
    j.npx = d.n_rows * d.n_cols;
    const bool prec = d.form == TDC_FORM_PRECISION;
    j.tile_px = tdc::kBlock * (prec ? tdc::kServePixelsPrecision : tdc::kServePixelsPlain);
    j.npt = (j.npx + j.tile_px - 1) / j.tile_px;
    const double per_chirp = static_cast<double>(std::min(j.npx, j.tile_px)) * d.n_samples;
    j.cps = std::max(1, static_cast<int>(kTileTerms / std::max(per_chirp, 1.0)));
    j.cps = std::min(j.cps, d.n_pri);
    j.nsplit = (d.n_pri + j.cps - 1) / j.cps;
    j.ntiles = static_cast<int64_t>(j.npt) * j.nsplit;
    j.nfin = (j.npx + tdc::kBlock - 1) / tdc::kBlock;
    j.terms = static_cast<double>(j.npx) * d.n_pri * d.n_samples;
    const int nblk = (d.n_samples + tdc::kAnchorBlock - 1) / tdc::kAnchorBlock;
    size_t b = 0;
    b += align256(8 * nterm);                                   // raw as uploaded
    b += align256(sizeof(tdc::ChirpPos) * d.n_pri);             // chirp positions
    b += align256(sizeof(float2) * j.npx);                      // image
    if (prec) {
        if (d.raw_storage != TDC_RAW_C64) {
            b += align256(4 * nterm);                           // converted raw
            b += align256(4 * static_cast<size_t>(d.n_samples)); // unit window for the conversion
        }
        b += align256(24 * static_cast<size_t>(j.npx));         // FP64 grid
        b += align256(8 * static_cast<size_t>(nblk));           // anchor table
        b += align256(sizeof(double2) * static_cast<size_t>(j.npx) * j.nsplit);
    } else {
        b += align256(4 * static_cast<size_t>(d.n_samples));    // k table
        b += align256(sizeof(tdc::PlainRec) * nterm);           // records
        b += align256(12 * static_cast<size_t>(j.npx));         // FP32 grid
        b += align256(sizeof(float2) * static_cast<size_t>(j.npx) * j.nsplit);
    }
    j.bytes = b;
}

int validate(const tdc_job_desc *d)
{
    if (!d) return sfail(TDC_ERR_INVALID_ARGUMENT, "job description is NULL");
    if (d->n_pri <= 0) return sfail(TDC_ERR_INVALID_ARGUMENT, "nPri must be greater than 0");
    if (d->n_samples <= 0) return sfail(TDC_ERR_INVALID_ARGUMENT, "nSamples must be greater than 0");
    if (d->n_rows <= 0 || d->n_cols <= 0)
        return sfail(TDC_ERR_INVALID_ARGUMENT, "the focus grid must not be empty");
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

// Upload and stage one admitted job (scheduler thread, compute stream).
bool upload_job(tdc_server *s, Job &j, std::string &err)
{
    const tdc_job_desc &d = j.d;
    const size_t nterm = static_cast<size_t>(d.n_pri) * d.n_samples;
    const bool prec = d.form == TDC_FORM_PRECISION;
    std::vector<double> kanchor;
    double dk = 0.0;
    if (prec && !tdc::precision_tables(d.sample_times, d.n_samples, d.start_freq, d.mod_rate,
                                       kanchor, dk, err))
        return false;
    // uploads and staging run on the copy stream, beside the step in flight;
    // the compute stream waits for them through an event
    hipStream_t cs = s->copy_stream;
    SRV_HIP(hipMallocAsync(&j.arena, j.bytes, cs));
    char *p = static_cast<char *>(j.arena);
    auto take = [&](size_t n) {
        char *q = p;
        p += align256(n);
        return static_cast<void *>(q);
    };
    void *raw = take(8 * nterm);
    auto *cpos = static_cast<tdc::ChirpPos *>(take(sizeof(tdc::ChirpPos) * d.n_pri));
    j.d_image = static_cast<float2 *>(take(sizeof(float2) * j.npx));
    std::vector<tdc::ChirpPos> hc(d.n_pri);
    for (int n = 0; n < d.n_pri; ++n) {
        const double *c = d.chirp_position + 3 * static_cast<size_t>(n);
        hc[n] = tdc::ChirpPos{c[0], c[1], c[2], 0.0};
    }
    SRV_HIP(hipMemcpyAsync(raw, d.raw_iq, 8 * nterm, hipMemcpyHostToDevice, cs));
    SRV_HIP(hipMemcpyAsync(cpos, hc.data(), sizeof(tdc::ChirpPos) * d.n_pri,
                           hipMemcpyHostToDevice, cs));
    tdc::ServeJob sj{};
    sj.cpos = cpos;
    sj.image = j.d_image;
    sj.npx = j.npx;
    sj.npri = d.n_pri;
    sj.ns = d.n_samples;
    sj.nsplit = j.nsplit;
    sj.cps = j.cps;
    sj.form = prec ? 1 : 0;
    sj.storage = d.raw_storage;
    std::vector<float> ones;
    if (prec) {
        const void *staged = raw;
        if (d.raw_storage != TDC_RAW_C64) {
            // served raw data arrive windowed: convert with a unit window
            void *conv = take(4 * nterm);
            auto *w = static_cast<float *>(take(4 * static_cast<size_t>(d.n_samples)));
            ones.assign(d.n_samples, 1.0F);
            SRV_HIP(hipMemcpyAsync(w, ones.data(), 4 * ones.size(), hipMemcpyHostToDevice, cs));
            SRV_HIP(tdc::launch_stage_precision(d.raw_storage, static_cast<const float *>(raw), w,
                                                d.n_pri, d.n_samples, conv, cs));
            staged = conv;
        }
        auto *grid = static_cast<double *>(take(24 * static_cast<size_t>(j.npx)));
        auto *ka = static_cast<double *>(take(8 * kanchor.size()));
        void *part = take(sizeof(double2) * static_cast<size_t>(j.npx) * j.nsplit);
        SRV_HIP(hipMemcpyAsync(grid, d.grid, 24 * static_cast<size_t>(j.npx),
                               hipMemcpyHostToDevice, cs));
        SRV_HIP(hipMemcpyAsync(ka, kanchor.data(), 8 * kanchor.size(), hipMemcpyHostToDevice, cs));
        sj.raw = staged;
        sj.grid64 = grid;
        sj.kanchor = ka;
        sj.dk = dk;
        sj.partial = part;
    } else {
        auto *kt = static_cast<float *>(take(4 * static_cast<size_t>(d.n_samples)));
        auto *rec = static_cast<tdc::PlainRec *>(take(sizeof(tdc::PlainRec) * nterm));
        auto *grid = static_cast<float *>(take(12 * static_cast<size_t>(j.npx)));
        void *part = take(sizeof(float2) * static_cast<size_t>(j.npx) * j.nsplit);
        std::vector<float> hk(d.n_samples), hg(3 * static_cast<size_t>(j.npx));
        for (int m = 0; m < d.n_samples; ++m)
            hk[m] = static_cast<float>(2.0 * (d.start_freq + d.mod_rate * d.sample_times[m]) / kC);
        for (size_t i = 0; i < hg.size(); ++i) hg[i] = static_cast<float>(d.grid[i]);
        SRV_HIP(hipMemcpyAsync(kt, hk.data(), 4 * hk.size(), hipMemcpyHostToDevice, cs));
        SRV_HIP(hipMemcpyAsync(grid, hg.data(), 4 * hg.size(), hipMemcpyHostToDevice, cs));
        SRV_HIP(tdc::launch_stage_plain_chirp(static_cast<const float *>(raw), cpos, kt, d.n_pri,
                                              d.n_samples, rec, cs));
        // the pageable host vectors above must outlive their copies
        SRV_HIP(hipStreamSynchronize(cs));
        sj.raw = rec;
        sj.grid32 = grid;
        sj.partial = part;
    }
    if (static_cast<size_t>(p - static_cast<char *>(j.arena)) > j.bytes) {
        err = "internal error: job footprint underestimated";
        return false;
    }
    SRV_HIP(hipMemcpyAsync(s->d_jobs + j.slot, &sj, sizeof sj, hipMemcpyHostToDevice, cs));
    SRV_HIP(hipEventRecord(s->ev_upload, cs));
    SRV_HIP(hipStreamWaitEvent(s->stream, s->ev_upload, 0));
    // the descriptor and the host-side tables are pageable: finish their copies
    SRV_HIP(hipStreamSynchronize(cs));
    return true;
}

bool ensure_graph(tdc_server *s, int bucket, int buf, std::string &err)
{
    if (s->graph[bucket][buf]) return true;
    hipStream_t cs = nullptr;
    SRV_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        e = tdc::launch_serve_step(s->d_jobs, s->d_table[buf], s->d_counter, kBuckets[bucket],
                                   kFinWorkgroups, cs);
        hipError_t e2 = hipStreamEndCapture(cs, &g);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&s->graph[bucket][buf], g, nullptr, nullptr, 0);
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

// Finish the jobs of a completed step (scheduler thread).
bool complete_step(tdc_server *s, Step &st, std::string &err)
{
    SRV_HIP(hipEventSynchronize(s->ev_end[st.buf]));
    float ms = 0.0F;
    SRV_HIP(hipEventElapsedTime(&ms, s->ev_start[st.buf], s->ev_end[st.buf]));
    for (Job *j : st.finishing) {
        j->image.resize(2 * static_cast<size_t>(j->npx));
        SRV_HIP(hipMemcpyAsync(j->image.data(), j->d_image, sizeof(float2) * j->npx,
                               hipMemcpyDeviceToHost, s->copy_stream));
    }
    SRV_HIP(hipStreamSynchronize(s->copy_stream));
    const double t = s->now();
    for (Job *j : st.finishing) SRV_HIP(hipFreeAsync(j->arena, s->copy_stream));
    std::lock_guard<std::mutex> lk(s->mu);
    s->busy_ms += ms;
    for (Job *j : st.finishing) {
        j->arena = nullptr;
        j->state = TDC_JOB_DONE;
        j->t_done = t;
        s->in_use -= j->bytes;
        s->free_slots.push_back(j->slot);
        s->terms_done += j->terms;
        ++s->completed;
        s->active.erase(std::find(s->active.begin(), s->active.end(), j));
    }
    s->cv_done.notify_all();
    return true;
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
        // 1. admission, in arrival order
        std::vector<Job *> admit;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->queued.empty() && s->active.empty() && inflight.empty()) {
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
                admit.push_back(j);
            }
        }
        for (Job *j : admit) {
            std::string why;
            if (!upload_job(s, *j, why)) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->in_use -= j->bytes;
                s->free_slots.push_back(j->slot);
                if (j->arena) (void) hipFreeAsync(j->arena, s->copy_stream);
                j->arena = nullptr;
                fail_job_locked(s, j, TDC_ERR_UNSUPPORTED, why);
                s->cv_done.notify_all();
                continue;
            }
            std::lock_guard<std::mutex> lk(s->mu);
            s->active.push_back(j);
        }

        // 2. the two-in-flight rule: finish the older step before reusing its buffer
        if (inflight.size() == 2) {
            if (!complete_step(s, inflight.front(), err)) return die(err);
            inflight.pop_front();
        }

        // 3. build the next step from the admitted jobs, oldest first
        Step st;
        st.index = step_index;
        st.buf = static_cast<int>(step_index & 1);
        tdc::StepTable *tab = s->h_table[st.buf];
        int nt = 0, nf = 0;
        double work = 0.0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            // processor sharing: every admitted job with tiles left gets an equal
            // share of the step (at least one tile), oldest first; what the
            // shares leave is then filled in arrival order
            int nwant = 0;
            for (Job *j : s->active)
                if (j->next_tile < j->ntiles) ++nwant;
            const double share = nwant > 0 ? s->step_terms / nwant : 0.0;
            auto take_tiles = [&](Job *j, double limit) {
                double got = 0.0;
                bool took = false;
                while (j->next_tile < j->ntiles && nt < tdc::kMaxStepTiles &&
                       (got < limit || !took) && (work < s->step_terms || nt == 0)) {
                    const int64_t t = j->next_tile++;
                    const int split = static_cast<int>(t / j->npt);
                    const int ptile = static_cast<int>(t % j->npt);
                    tab->tiles[nt++] = tdc::ServeTile{j->slot, ptile * j->tile_px, split, 0};
                    const int n0 = split * j->cps;
                    const int nch = std::min(j->d.n_pri, n0 + j->cps) - n0;
                    const int npx = std::min(j->tile_px, j->npx - ptile * j->tile_px);
                    const double w = static_cast<double>(npx) * nch * j->d.n_samples;
                    work += w;
                    got += w;
                    took = true;
                }
                if (took) {
                    if (j->t_first < 0) j->t_first = s->now();
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
                    for (int f = 0; f < j->nfin; ++f) tab->fin[nf++] = tdc::FinTile{j->slot, f * tdc::kBlock};
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
        tab->ntiles = nt;
        tab->nfin = nf;
        int bucket = kNumBuckets - 1;
        for (int b = 0; b < kNumBuckets; ++b) {
            if (kBuckets[b] >= nt) {
                bucket = b;
                break;
            }
        }
        if (!ensure_graph(s, bucket, st.buf, err)) return die(err);
        const size_t bytes = offsetof(tdc::StepTable, tiles) + sizeof(tdc::ServeTile) * nt;
        hipError_t e = hipMemcpyAsync(s->d_table[st.buf], tab, bytes, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess && nf > 0)
            e = hipMemcpyAsync(&s->d_table[st.buf]->fin[0], &tab->fin[0], sizeof(tdc::FinTile) * nf,
                               hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = hipEventRecord(s->ev_start[st.buf], s->stream);
        if (e == hipSuccess) e = hipGraphLaunch(s->graph[bucket][st.buf], s->stream);
        if (e == hipSuccess) e = hipEventRecord(s->ev_end[st.buf], s->stream);
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
        SRV_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        SRV_HIP(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
        for (int i = 0; i < 2; ++i) {
            SRV_HIP(hipEventCreate(&s->ev_start[i]));
            SRV_HIP(hipEventCreate(&s->ev_end[i]));
            SRV_HIP(hipMalloc(reinterpret_cast<void **>(&s->d_table[i]), sizeof(tdc::StepTable)));
            SRV_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->h_table[i]), sizeof(tdc::StepTable),
                                  hipHostMallocDefault));
        }
        SRV_HIP(hipMalloc(reinterpret_cast<void **>(&s->d_jobs), sizeof(tdc::ServeJob) * tdc::kMaxJobSlots));
        SRV_HIP(hipMalloc(reinterpret_cast<void **>(&s->d_counter), 256));
        SRV_HIP(hipEventCreateWithFlags(&s->ev_upload, hipEventDisableTiming));
        return true;
    };
    if (!init()) {
        tdc_server_destroy(s.release());
        return sfail(TDC_ERR_DEVICE, err);
    }
    for (int i = tdc::kMaxJobSlots - 1; i >= 0; --i) s->free_slots.push_back(i);
    tdc_server *raw = s.release();
    raw->worker = std::thread(scheduler, raw);
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
    if (image_iq) std::memcpy(image_iq, j->image.data(), sizeof(float) * j->image.size());
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
    (void) hipSetDevice(s->device);
    if (s->stream) (void) hipStreamSynchronize(s->stream);
    for (auto &row : s->graph)
        for (auto &g : row)
            if (g) (void) hipGraphExecDestroy(g);
    for (int i = 0; i < 2; ++i) {
        if (s->ev_start[i]) (void) hipEventDestroy(s->ev_start[i]);
        if (s->ev_end[i]) (void) hipEventDestroy(s->ev_end[i]);
        if (s->d_table[i]) (void) hipFree(s->d_table[i]);
        if (s->h_table[i]) (void) hipHostFree(s->h_table[i]);
    }
    if (s->ev_upload) (void) hipEventDestroy(s->ev_upload);
    if (s->d_jobs) (void) hipFree(s->d_jobs);
    if (s->d_counter) (void) hipFree(s->d_counter);
    if (s->copy_stream) (void) hipStreamDestroy(s->copy_stream);
    if (s->stream) (void) hipStreamDestroy(s->stream);
    delete s;
    return TDC_OK;
}

} // extern "C"
