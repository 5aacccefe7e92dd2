// tdc_serve.hip -- kernels of the serving form.
//
// A step of the server runs one persistent kernel per tile family it holds,
// captured once per grid-size bucket in a hipGraph and replayed, and a
// finalize kernel behind it:
//   1. serve_kernel         persistent: every workgroup takes the next tile
//                           from the step table with one atomic on a counter
//                           that is never reset (the step table holds the
//                           counter's value at the step's start), runs the
//                           tile body of its job (plain; precision with the
//                           range-profile or matrix engine; any raw storage)
//                           and writes the job's partial image for that
//                           (pixel tile, chirp range); all workgroups leave
//                           once the counter passes the tile count;
//   2. serve_finalize_kernel sums each finished job's partial images in
//                           chirp-range order into its image.
// A job's tiles, their chirp ranges and the finalize order depend only on
// the job's own shape, so its image is bit-identical whatever else shares the
// step, and whichever workgroup runs which tile.
#include <hip/hip_runtime.h>

#include "tdc_serve.h"
#include "tdc_tiles.h"

namespace tdc {
namespace {

// A pointer that is the same in every lane, as the compiler's uniform value.
template <class T> __device__ __forceinline__ const T *uniform_ptr(const T *p)
{
    const unsigned long long a = reinterpret_cast<unsigned long long>(p);
    const unsigned int lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(a));
    const unsigned int hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(a >> 32));
    return reinterpret_cast<const T *>((static_cast<unsigned long long>(hi) << 32) | lo);
}

struct ServeTileIndex {
    int pix0;  // first pixel of the tile
    int split; // chirp range index
};

// FAMILY 0 runs the table's first nfast tiles (plain form and range-profile
// engine, at most ~62 VGPRs, 8 waves per SIMD), FAMILY 1 the rest (matrix
// engine, which needs ~130 VGPRs); each family has its own kernel so that
// the matrix tile's registers do not limit the occupancy of the others.
template <int FAMILY>
__global__ __launch_bounds__(kBlock) void serve_kernel(const StepTable *__restrict__ table,
                                                       unsigned int *__restrict__ counter)
{
    __shared__ int s_first[kMaxStepRuns];
    __shared__ int s_tile0[kMaxStepRuns];
    __shared__ const ServeJob *s_job[kMaxStepRuns];
    __shared__ unsigned int s_next;
    const cptr<StepTable> tab = as_const(table);
    const unsigned int count = static_cast<unsigned int>(tab->ntiles[FAMILY]);
    const int r0 = FAMILY == 0 ? 0 : tab->nruns[0];
    const int nruns = tab->nruns[FAMILY];
    const unsigned int base = tab->base[FAMILY];
    // the family's runs, read once from the step table
    for (int r = static_cast<int>(threadIdx.x); r < nruns; r += kBlock) {
        const StepRun run = table->runs[r0 + r];
        s_first[r] = run.first;
        s_tile0[r] = table->tile0[r0 + r];
        s_job[r] = run.job;
    }
    for (;;) {
        // the next tile (taking it when the previous one has ended keeps the
        // tiles balanced over the workgroups to the last one)
        if (threadIdx.x == 0) s_next = atomicAdd(counter + FAMILY, 1U) - base;
        __syncthreads();
        const unsigned int t = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();
        if (t >= count) break; // every wave of the workgroup takes this exit
        // the run holding tile t: the last run whose first tile is <= t
        int lo = 0, hi = nruns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (static_cast<unsigned int>(s_first[mid]) <= t) lo = mid;
            else hi = mid - 1;
        }
        // values read from LDS are uniform here: made so for the compiler, which
        // then keeps the job's descriptor and every loop bound in SGPRs
        lo = __builtin_amdgcn_readfirstlane(lo);
        const int jt = __builtin_amdgcn_readfirstlane(s_tile0[lo]) + static_cast<int>(t) -
                       __builtin_amdgcn_readfirstlane(s_first[lo]);
        const ServeJob j = as_const(uniform_ptr(s_job[lo]))[0];
        const int split = jt / j.npt;
        const int pix0 = (jt - split * j.npt) * kServeTilePixels;
        const ServeTileIndex tl{pix0, split};
        const int n0 = tl.split * j.cps;
        const int n1 = min(j.npri, n0 + j.cps);
        if (FAMILY == 0) {
            if (j.form == 0) {
                range_tile<1, double, ChirpPos>(static_cast<const float4 *>(j.raw), j.cpos,
                                                j.grid64, j.kanchor, j.npx, j.ns, n0, n1,
                                                tl.pix0,
                                                static_cast<float2 *>(j.partial) +
                                                    static_cast<size_t>(tl.split) * j.npx);
            } else {
                const auto *tab = static_cast<const float2 *>(j.raw);
                float2 *out = static_cast<float2 *>(j.partial) +
                              static_cast<size_t>(tl.split) * j.npx;
                // served jobs run the automatic kernel: (6, 4), or (10, 2) for small
                // images and chirps whose fourfold transform would not fit
                if (j.pp.over == 4)
                    profile_tile<6, 4>(tab, j.cpos, j.grid64, j.pp, j.npx, n0, n1, tl.pix0, out);
                else
                    profile_tile<10, 2>(tab, j.cpos, j.grid64, j.pp, j.npx, n0, n1, tl.pix0, out);
            }
        } else {
            double2 *out = static_cast<double2 *>(j.partial) + static_cast<size_t>(tl.split) * j.npx;
            const auto *atab = static_cast<const uintx4_t *>(j.raw);
            const int nq = matrix_quads(j.ns);
            if (j.storage == 0) {
                matrix_tile<2>(atab, j.inv_scale, j.cpos, j.grid64, j.kanchor, j.dk, j.npx, nq, n0,
                               n1, tl.pix0, out);
            } else {
                matrix_tile<1>(atab, j.inv_scale, j.cpos, j.grid64, j.kanchor, j.dk, j.npx, nq, n0,
                               n1, tl.pix0, out);
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void serve_finalize_kernel(const StepTable *__restrict__ table)
{
    const cptr<StepTable> tab = as_const(table);
    const int nfin = tab->nfin;
    for (int f = static_cast<int>(blockIdx.x); f < nfin; f += static_cast<int>(gridDim.x)) {
        const FinTile ft = tab->fin[f];
        const ServeJob j = as_const(ft.job)[0];
        const int p = ft.pix0 + static_cast<int>(threadIdx.x);
        if (p >= j.npx) continue;
        if (j.form == 0) {
            const float2 *part = static_cast<const float2 *>(j.partial);
            float sr = 0.0F, si = 0.0F;
            // unrolled so that the loads of several splits are in flight at
            // once; the adds stay in split order
#pragma unroll 8
            for (int s = 0; s < j.nsplit; ++s) {
                const float2 v = part[static_cast<size_t>(s) * j.npx + p];
                sr += v.x;
                si += v.y;
            }
            j.image[p] = make_float2(sr, si);
        } else if (j.profile) {
            // FP32 partial images of the range-profile tiles, added in FP64
            const float2 *part = static_cast<const float2 *>(j.partial);
            double sr = 0.0, si = 0.0;
#pragma unroll 8
            for (int s = 0; s < j.nsplit; ++s) {
                const float2 v = part[static_cast<size_t>(s) * j.npx + p];
                sr += static_cast<double>(v.x);
                si += static_cast<double>(v.y);
            }
            j.image[p] = make_float2(static_cast<float>(sr), static_cast<float>(si));
        } else {
            const double2 *part = static_cast<const double2 *>(j.partial);
            double sr = 0.0, si = 0.0;
#pragma unroll 8
            for (int s = 0; s < j.nsplit; ++s) {
                const double2 v = part[static_cast<size_t>(s) * j.npx + p];
                sr += v.x;
                si += v.y;
            }
            j.image[p] = make_float2(static_cast<float>(sr), static_cast<float>(si));
        }
    }
}

// Plain-form records of a served job (stop-and-hop: the FP64 phase centre of
// each chirp is read from the job's chirp table): terms (rr, ri, dk_m, 0).
__global__ __launch_bounds__(kBlock) void stage_plain_chirp_kernel(const float2 *__restrict__ raw,
                                                                   const float *__restrict__ dk,
                                                                   int npri, int ns,
                                                                   float4 *__restrict__ terms)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    const size_t stride = static_cast<size_t>(gridDim.x) * kBlock;
    const size_t first = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
    for (size_t e = first; e < total; e += stride) {
        const int m = static_cast<int>(e % ns);
        const float2 v = raw[e];
        terms[e] = make_float4(v.x, v.y, dk[m], 0.0F);
    }
}

} // namespace

hipError_t launch_serve_step(const StepTable *table, unsigned int *counter, int workgroups,
                             bool matrix, hipStream_t s)
{
    hipLaunchKernelGGL(serve_kernel<0>, dim3(workgroups), dim3(kBlock), 0, s, table, counter);
    if (matrix)
        hipLaunchKernelGGL(serve_kernel<1>, dim3(workgroups), dim3(kBlock), 0, s, table, counter);
    return hipGetLastError();
}

hipError_t launch_serve_finalize(const StepTable *table, int workgroups, hipStream_t s)
{
    hipLaunchKernelGGL(serve_finalize_kernel, dim3(workgroups), dim3(kBlock), 0, s, table);
    return hipGetLastError();
}

hipError_t launch_stage_plain_chirp(const float *raw, const float *dk, int npri, int ns,
                                    float4 *terms, hipStream_t s)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    const size_t want = (total + kBlock - 1) / kBlock;
    const int grid = static_cast<int>(want < 4096 ? (want > 0 ? want : 1) : 4096);
    hipLaunchKernelGGL(stage_plain_chirp_kernel, dim3(grid), dim3(kBlock), 0, s,
                       reinterpret_cast<const float2 *>(raw), dk, npri, ns, terms);
    return hipGetLastError();
}

} // namespace tdc
