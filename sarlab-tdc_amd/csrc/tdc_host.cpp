// tdc_host.cpp -- host-side ingest of one image's inputs.
//
// The reference stages each 128-chirp chunk through page-locked buffers with a
// blocking copy per chunk and pixel (tdc-lib/src/tdcprocessor.cpp:320-324,
// gpuarray.h:79-87); its asynchronous copy helpers and stream wrapper
// (gpuarray.h:95-103, cudastream.h:17-64) are never used. Here the inputs of
// one image are checked by a fused pass that a small pool of host threads
// splits by chirp, and copied to the device through two page-locked slots on
// a copy stream of their own: the pool fills one slot while the DMA engine
// drains the other, and the compute stream waits on one event.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

#include <pthread.h>

#include <hip/hip_runtime.h>

#include "tdc_internal.h"

namespace tdc {

namespace {

// Worker threads that run contiguous parts of a range. The calling thread runs
// one part itself; calls from several threads are serialised.
class Pool {
  public:
    explicit Pool(int threads)
    {
        for (int i = 0; i + 1 < threads; ++i) workers_.emplace_back([this] { work(); });
    }
    // The threads live until the process ends (no join at exit).
    ~Pool() = delete;

    int size() const { return static_cast<int>(workers_.size()) + 1; }

    void run(int parts, const std::function<void(int)> &fn)
    {
        std::lock_guard<std::mutex> serial(run_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            parts_ = parts;
            next_.store(0);
            left_ = parts;
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        drain(fn, parts);
        std::unique_lock<std::mutex> lk(mu_);
        // no worker may still hold `fn` when it goes out of scope
        done_cv_.wait(lk, [this] { return left_ == 0 && active_ == 0; });
        fn_ = nullptr;
    }

  private:
    void drain(const std::function<void(int)> &fn, int parts)
    {
        for (;;) {
            const int i = next_.fetch_add(1);
            if (i >= parts) return;
            fn(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_cv_.notify_all();
        }
    }

    void work()
    {
        unsigned long seen = 0;
        for (;;) {
            // calls come in bursts (one per copied piece of an upload): poll
            // for a while before sleeping, so the next call finds the
            // workers awake
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load() == seen &&
                   std::chrono::steady_clock::now() - t0 < std::chrono::microseconds(kSpinMicros))
                __builtin_ia32_pause();
            const std::function<void(int)> *fn;
            int parts;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_.load() != seen; });
                seen = gen_.load();
                if (!fn_) continue; // woke after that call had ended
                fn = fn_;
                parts = parts_;
                ++active_;
            }
            drain(*fn, parts);
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0) done_cv_.notify_all();
        }
    }

    std::vector<std::thread> workers_;
    std::mutex run_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *fn_ = nullptr;
    int parts_ = 0, left_ = 0, active_ = 0;
    std::atomic<int> next_{0};
    std::atomic<unsigned long> gen_{0};
    static constexpr int kSpinMicros = 300;
};

// A child process created by fork() has none of the parent's worker threads:
// it starts a pool of its own (the parent's object is left alone).
std::atomic<Pool *> g_pool{nullptr};
std::mutex g_pool_mu;

void forget_pool_in_child() { g_pool.store(nullptr); }

Pool &pool()
{
    Pool *p = g_pool.load();
    if (p) return *p;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    p = g_pool.load();
    if (!p) {
        static std::once_flag hook;
        std::call_once(hook, [] { pthread_atfork(nullptr, nullptr, forget_pool_in_child); });
        const unsigned hw = std::thread::hardware_concurrency();
        p = new Pool(static_cast<int>(std::max(1U, std::min(kHostThreads, hw))));
        g_pool.store(p);
    }
    return *p;
}

} // namespace

void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)> &fn)
{
    if (n == 0) return;
    grain = std::max<size_t>(grain, 1);
    Pool &p = pool();
    const size_t parts = std::min<size_t>(static_cast<size_t>(p.size()), (n + grain - 1) / grain);
    if (parts <= 1) {
        fn(0, n);
        return;
    }
    const size_t step = (n + parts - 1) / parts;
    p.run(static_cast<int>(parts), [&](int i) {
        const size_t b = static_cast<size_t>(i) * step;
        if (b < n) fn(b, std::min(n, b + step));
    });
}

void grid_bounds(const double *grid, size_t npx, double lo[3], double hi[3])
{
    const double inf = std::numeric_limits<double>::infinity();
    for (int a = 0; a < 3; ++a) {
        lo[a] = inf;
        hi[a] = -inf;
    }
    std::mutex mu;
    bool nan = false;
    parallel_for(npx, size_t(1) << 15, [&](size_t i0, size_t i1) {
        double l[3] = {inf, inf, inf}, h[3] = {-inf, -inf, -inf};
        bool bad = false;
        for (size_t i = i0; i < i1; ++i) {
            for (int a = 0; a < 3; ++a) {
                const double v = grid[3 * i + a];
                bad = bad || v != v;
                l[a] = std::min(l[a], v);
                h[a] = std::max(h[a], v);
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        nan = nan || bad;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    });
    if (nan)
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = std::numeric_limits<double>::quiet_NaN();
}

namespace {

inline uint32_t bits_of(float v)
{
    uint32_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

// 1 when the float with bit pattern b is an infinity or a NaN
inline uint32_t nonfinite(uint32_t b) { return (b & 0x7F800000U) == 0x7F800000U ? 1U : 0U; }

// The samples of chirps [n0, n1): bit patterns are compared with the chirp's
// first sample by XOR and OR-accumulated, exponents tested for all ones, and
// no branch depends on the data, so that the inner loops vectorise.
void scan_chirps(const float *pos, const float *vel, const float *att, int ns, size_t n0,
                 size_t n1, AntennaScan &s)
{
    uint32_t bad = 0, dpos = 0, dvel = 0, datt = 0;
    float lo0 = s.lo[0], lo1 = s.lo[1], lo2 = s.lo[2], hi0 = s.hi[0], hi1 = s.hi[1], hi2 = s.hi[2];
    double vmax2 = s.vmax2, qmax2 = s.qmax2, dmax2 = s.dmax2;
    for (size_t n = n0; n < n1; ++n) {
        const size_t e0 = n * static_cast<size_t>(ns);
        const float *p = pos + 3 * e0;
        const uint32_t px = bits_of(p[0]), py = bits_of(p[1]), pz = bits_of(p[2]);
        for (int m = 0; m < ns; ++m) {
            const float x = p[3 * m], y = p[3 * m + 1], z = p[3 * m + 2];
            const uint32_t bx = bits_of(x), by = bits_of(y), bz = bits_of(z);
            bad |= nonfinite(bx) | nonfinite(by) | nonfinite(bz);
            dpos |= (bx ^ px) | (by ^ py) | (bz ^ pz);
            const double ox = double(x) - double(p[0]), oy = double(y) - double(p[1]),
                         oz = double(z) - double(p[2]);
            dmax2 = std::max(dmax2, ox * ox + oy * oy + oz * oz);
            lo0 = std::min(lo0, x);
            lo1 = std::min(lo1, y);
            lo2 = std::min(lo2, z);
            hi0 = std::max(hi0, x);
            hi1 = std::max(hi1, y);
            hi2 = std::max(hi2, z);
        }
        if (vel) {
            const float *v = vel + 3 * e0;
            const uint32_t vx = bits_of(v[0]), vy = bits_of(v[1]), vz = bits_of(v[2]);
            for (int m = 0; m < ns; ++m) {
                const float x = v[3 * m], y = v[3 * m + 1], z = v[3 * m + 2];
                const uint32_t bx = bits_of(x), by = bits_of(y), bz = bits_of(z);
                bad |= nonfinite(bx) | nonfinite(by) | nonfinite(bz);
                dvel |= (bx ^ vx) | (by ^ vy) | (bz ^ vz);
                vmax2 = std::max(vmax2, double(x) * x + double(y) * y + double(z) * z);
            }
        }
        if (att) {
            const float *q = att + 4 * e0;
            const uint32_t qa = bits_of(q[0]), qb = bits_of(q[1]), qc = bits_of(q[2]),
                           qd = bits_of(q[3]);
            for (int m = 0; m < ns; ++m) {
                const float a = q[4 * m], b = q[4 * m + 1], c = q[4 * m + 2], d = q[4 * m + 3];
                const uint32_t ba = bits_of(a), bb = bits_of(b), bc = bits_of(c), bd = bits_of(d);
                bad |= nonfinite(ba) | nonfinite(bb) | nonfinite(bc) | nonfinite(bd);
                datt |= (ba ^ qa) | (bb ^ qb) | (bc ^ qc) | (bd ^ qd);
                qmax2 = std::max(qmax2, double(a) * a + double(b) * b + double(c) * c +
                                            double(d) * d);
            }
        }
    }
    s.dmax2 = dmax2;
    s.finite = s.finite && bad == 0;
    s.pos_const = s.pos_const && dpos == 0;
    s.vel_const = s.vel_const && dvel == 0;
    s.att_const = s.att_const && datt == 0;
    s.lo[0] = lo0;
    s.lo[1] = lo1;
    s.lo[2] = lo2;
    s.hi[0] = hi0;
    s.hi[1] = hi1;
    s.hi[2] = hi2;
    s.vmax2 = vmax2;
    s.qmax2 = qmax2;
}

} // namespace

AntennaScan scan_antenna(const float *pos, const float *vel, const float *att, int npri, int ns)
{
    AntennaScan out;
    out.has_vel = vel != nullptr;
    out.has_att = att != nullptr;
    if (!pos || npri <= 0 || ns <= 0) return out;
    for (int a = 0; a < 3; ++a) out.lo[a] = out.hi[a] = pos[a];
    const AntennaScan start = out;
    std::mutex mu;
    // ~16 KiB of positions per part at the least
    const size_t grain = std::max<size_t>(1, 1400 / static_cast<size_t>(ns));
    parallel_for(static_cast<size_t>(npri), grain, [&](size_t n0, size_t n1) {
        AntennaScan part = start;
        scan_chirps(pos, vel, att, ns, n0, n1, part);
        std::lock_guard<std::mutex> lk(mu);
        out.finite = out.finite && part.finite;
        out.pos_const = out.pos_const && part.pos_const;
        out.vel_const = out.vel_const && part.vel_const;
        out.att_const = out.att_const && part.att_const;
        out.vmax2 = std::max(out.vmax2, part.vmax2);
        out.qmax2 = std::max(out.qmax2, part.qmax2);
        out.dmax2 = std::max(out.dmax2, part.dmax2);
        for (int a = 0; a < 3; ++a) {
            out.lo[a] = std::min(out.lo[a], part.lo[a]);
            out.hi[a] = std::max(out.hi[a], part.hi[a]);
        }
    });
    return out;
}

// ---------------------------------------------------------------------------
// Ingest: page-locked double buffering on a copy stream.
// ---------------------------------------------------------------------------

Ingest::~Ingest()
{
    if (copy_) (void) hipStreamSynchronize(copy_);
    for (int i = 0; i < 2; ++i) {
        if (free_[i]) (void) hipEventDestroy(free_[i]);
        if (slot_[i]) (void) hipHostFree(slot_[i]);
    }
    if (tail_) (void) hipEventDestroy(tail_);
    if (copy_) (void) hipStreamDestroy(copy_);
}

// Creates what is still missing (a failed call leaves the rest for the next).
hipError_t Ingest::init()
{
    if (ready_) return hipSuccess;
    hipError_t e = hipSuccess;
    if (!copy_ && (e = hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking)) != hipSuccess) {
        copy_ = nullptr;
        return e;
    }
    for (int i = 0; i < 2; ++i) {
        if (!slot_[i] && (e = hipHostMalloc(&slot_[i], kSlotBytes, hipHostMallocDefault)) != hipSuccess) {
            slot_[i] = nullptr;
            return e;
        }
        if (!free_[i] && (e = hipEventCreateWithFlags(&free_[i], hipEventDisableTiming)) != hipSuccess) {
            free_[i] = nullptr;
            return e;
        }
    }
    if (!tail_ && (e = hipEventCreateWithFlags(&tail_, hipEventDisableTiming)) != hipSuccess) {
        tail_ = nullptr;
        return e;
    }
    ready_ = true;
    return hipSuccess;
}

hipError_t Ingest::begin(hipStream_t compute)
{
    hipError_t e = init();
    if (e != hipSuccess) return e;
    // device buffers being refilled may still be read by queued compute work
    if ((e = hipEventRecord(tail_, compute)) != hipSuccess) return e;
    return hipStreamWaitEvent(copy_, tail_, 0);
}

hipError_t Ingest::put(void *dst, const void *src, size_t bytes)
{
    hipError_t e = init();
    if (e != hipSuccess) return e;
    const char *from = static_cast<const char *>(src);
    char *to = static_cast<char *>(dst);
    while (bytes > 0) {
        const size_t piece = std::min(bytes, kPieceBytes);
        if (used_ + piece > kSlotBytes) {
            // the other slot is free once the copies issued from it have ended
            Range wait("tdc_ingest_slot_wait");
            if ((e = hipEventRecord(free_[cur_], copy_)) != hipSuccess) return e;
            cur_ ^= 1;
            if ((e = hipEventSynchronize(free_[cur_])) != hipSuccess) return e;
            used_ = 0;
        }
        char *slot = static_cast<char *>(slot_[cur_]) + used_;
        {
            Range fill("tdc_ingest_slot_fill");
            if (piece >= kParallelCopyBytes) {
                parallel_for(piece, kParallelCopyBytes / 2, [&](size_t b, size_t end) {
                    std::memcpy(slot + b, from + b, end - b);
                });
            } else {
                std::memcpy(slot, from, piece);
            }
        }
        if ((e = hipMemcpyAsync(to, slot, piece, hipMemcpyHostToDevice, copy_)) != hipSuccess)
            return e;
        used_ += (piece + 255) & ~static_cast<size_t>(255);
        from += piece;
        to += piece;
        bytes -= piece;
    }
    return hipSuccess;
}

hipError_t Ingest::wait(hipEvent_t done)
{
    hipError_t e = init();
    if (e != hipSuccess) return e;
    return hipStreamWaitEvent(copy_, done, 0);
}

hipError_t Ingest::fence(hipStream_t compute)
{
    hipError_t e = init();
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(tail_, copy_)) != hipSuccess) return e;
    return hipStreamWaitEvent(compute, tail_, 0);
}

} // namespace tdc
