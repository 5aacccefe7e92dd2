// tdc_kernels.hip -- gfx950 kernels of the TDC correlate-sum path.
//
// Reference path (tdc-lib/src/tdckernels.cu): for every pixel and every
// 128-chirp chunk the reference launches createWindowKernel (:70-132),
// referenceResponseKernel (:178-217), correlateWithReference (:250-274), an
// NPP complex sum and addToImage (:276-279, :309-314), each materialising a
// chunk-sized array in global memory. Here one launch forms the whole image:
//
//   * every lane owns P pixels (coordinates and accumulators in VGPRs);
//   * the loop runs over (chirp, sample) terms, which are the same for every
//     lane of a wavefront, so each term's record arrives through one scalar
//     load into SGPRs and is consumed as the scalar operand of the VALU ops;
//   * chirps are split into ranges across workgroups when the grid alone
//     cannot fill the chip; the partial images are summed in a fixed order
//     by a second small kernel, so results do not depend on scheduling.
//
// Plain form (the drop-in): per term the direct reference phase
//   theta = k_m * |pos - p|  (cycles, k_m = 2 (f0 + K t_m) / c)
// from an FP64 range and an FP64 anchor phase every 64 samples plus one FP32
// FMA per term, with hardware fract / sin / cos (inputs in revolutions), then
// a complex multiply-accumulate of raw * exp(+j 2 pi theta) -- the conjugate
// product of tdckernels.cu:270-272 with ref = w exp(-j 4 pi f R / c).
//
// Precision form: per (pixel, chirp) the range R is computed in FP64; per
// 64-sample block the anchor phase k_a R is reduced modulo one cycle in
// FP64; inside the block the phase advances by the constant rotation
// w = exp(j 2 pi dk R), applied in nested (Horner) form to the running block
// sum, S <- S w + raw, which is one complex rotate-and-add (4 FMAs) per term.
// Block sums are rotated by the anchor phasor and accumulated with an FP32
// two-sum compensation.
#include <hip/hip_runtime.h>

#include "tdc_internal.h"
#include "tdc_tiles.h"

namespace tdc {
namespace {

// Workgroup -> (pixel tile, chirp split). Workgroups are dispatched to the
// eight XCDs round-robin by id, so with xcd_map the workgroups of one XCD
// share a chirp range and that range's records stay in the XCD's own L2.
__device__ __forceinline__ void map_block(const Geom &g, int &tile, int &split)
{
    const int b = static_cast<int>(blockIdx.x);
    if (g.xcd_map) {
        const int xcd = b & 7;
        const int slot = b >> 3;
        if (g.nsplit >= 8) {
            const int r = g.nsplit >> 3;
            split = xcd + 8 * (slot % r);
            tile = slot / r;
        } else {
            const int per = 8 / g.nsplit;
            split = xcd % g.nsplit;
            tile = slot * per + xcd / g.nsplit;
        }
    } else {
        split = b % g.nsplit;
        tile = b / g.nsplit;
    }
}

// Quaternion product / rotation in FP32 (reference tdc-lib/include/gpumath.h:66-87).
__device__ __forceinline__ float4 q_prod(float4 a, float4 b)
{
    return make_float4((a.x * b.x) - (a.y * b.y) - (a.z * b.z) - (a.w * b.w),
                       (a.x * b.y) + (a.y * b.x) + (a.z * b.w) - (a.w * b.z),
                       (a.x * b.z) - (a.y * b.w) + (a.z * b.x) + (a.w * b.y),
                       (a.x * b.w) + (a.y * b.z) - (a.z * b.y) + (a.w * b.x));
}

__device__ __forceinline__ float3 q_rot(float4 q, float3 v)
{
    const float4 pv = make_float4(0.0F, v.x, v.y, v.z);
    const float4 qc = make_float4(q.x, -q.y, -q.z, -q.w);
    const float4 r = q_prod(q_prod(q, pv), qc);
    return make_float3(r.y, r.z, r.w);
}

// Doppler window centre of one term: the centroid (2/lambda_c) v . a with the
// antenna pointing a = q b q* (windowing.h:19-24, gpumath.h:77-87), or the
// fixed centre.
__device__ __forceinline__ float window_centre(float3 vv, float4 att, float3 bore,
                                               float two_over_lambda, int centroid, float fc_fixed)
{
    if (!centroid) return fc_fixed;
    const float3 ant = q_rot(att, bore);
    return two_over_lambda * ((vv.x * ant.x) + (vv.y * ant.y) + (vv.z * ant.z));
}

// ---------------------------------------------------------------------------
// Staging: raw * range window, antenna offsets and dk_m packed per term.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void stage_plain_kernel(
    const float2 *__restrict__ raw, const float *__restrict__ pos, const float *__restrict__ dk,
    const float *__restrict__ wr, int npri, int ns, PlainRec *__restrict__ out)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    for (size_t e = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * kBlock) {
        const int m = static_cast<int>(e % ns);
        const size_t e0 = e - static_cast<size_t>(m); // the chirp's first sample
        const float2 v = raw[e];
        PlainRec r;
        r.rr = v.x * wr[m];
        r.ri = v.y * wr[m];
        // offset from the chirp's first position (exactly zero on stop-and-hop input)
        r.x = pos[3 * e] - pos[3 * e0];
        r.y = pos[3 * e + 1] - pos[3 * e0 + 1];
        r.z = pos[3 * e + 2] - pos[3 * e0 + 2];
        r.k = dk[m];
        r.pad0 = 0.0F;
        r.pad1 = 0.0F;
        out[e] = r;
    }
}

__global__ __launch_bounds__(kBlock) void stage_doppler_kernel(
    const float2 *__restrict__ raw, const float *__restrict__ pos, const float *__restrict__ vel,
    const float4 *__restrict__ att, const float *__restrict__ dk, const float *__restrict__ wr,
    int npri, int ns, float two_over_lambda, float3 bore, int centroid, float fc_fixed,
    DopRec *__restrict__ out)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    for (size_t e = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * kBlock) {
        const int m = static_cast<int>(e % ns);
        const float2 v = raw[e];
        const float3 vv = make_float3(vel[3 * e], vel[3 * e + 1], vel[3 * e + 2]);
        DopRec r;
        r.rr = v.x * wr[m];
        r.ri = v.y * wr[m];
        r.x = pos[3 * e];
        r.y = pos[3 * e + 1];
        r.z = pos[3 * e + 2];
        r.k = dk[m];
        r.vx = two_over_lambda * vv.x;
        r.vy = two_over_lambda * vv.y;
        r.vz = two_over_lambda * vv.z;
        r.fc = window_centre(vv, att[e], bore, two_over_lambda, centroid, fc_fixed);
        r.pad0 = 0.0F;
        r.pad1 = 0.0F;
        out[e] = r;
    }
}

// Stop-and-hop terms: (raw * range window, dk_m), the same products as
// stage_plain_kernel.
__global__ __launch_bounds__(kBlock) void stage_terms_kernel(const float2 *__restrict__ raw,
                                                             const float *__restrict__ dk,
                                                             const float *__restrict__ wr,
                                                             int npri, int ns,
                                                             float4 *__restrict__ out)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    for (size_t e = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * kBlock) {
        const int m = static_cast<int>(e % ns);
        const float2 v = raw[e];
        out[e] = make_float4(v.x * wr[m], v.y * wr[m], dk[m], 0.0F);
    }
}

// Per chirp: the position of its first sample and, when `vel` is given, the
// velocity and window centre of stage_doppler_kernel for that sample.
__global__ __launch_bounds__(kBlock) void stage_chirps_kernel(
    const float *__restrict__ pos, const float *__restrict__ vel, const float4 *__restrict__ att,
    int npri, int ns, float two_over_lambda, float3 bore, int centroid, float fc_fixed,
    float4 *__restrict__ cpos, float4 *__restrict__ cvel)
{
    for (int n = static_cast<int>(blockIdx.x * kBlock + threadIdx.x); n < npri;
         n += static_cast<int>(gridDim.x) * kBlock) {
        const size_t e = static_cast<size_t>(n) * ns;
        cpos[n] = make_float4(pos[3 * e], pos[3 * e + 1], pos[3 * e + 2], 0.0F);
        if (vel) {
            const float3 vv = make_float3(vel[3 * e], vel[3 * e + 1], vel[3 * e + 2]);
            cvel[n] = make_float4(two_over_lambda * vv.x, two_over_lambda * vv.y,
                                  two_over_lambda * vv.z,
                                  window_centre(vv, att[e], bore, two_over_lambda, centroid,
                                                fc_fixed));
        }
    }
}

// ---------------------------------------------------------------------------
// Plain form, Doppler window identically one (proved on the host).
// ---------------------------------------------------------------------------

template <int P, bool SERIES>
__global__ __launch_bounds__(kBlock) void plain_kernel(const PlainRec *__restrict__ rec,
                                                       const float4 *__restrict__ cpos,
                                                       const float *__restrict__ grid,
                                                       const double *__restrict__ kanchor, Geom g,
                                                       float2 *__restrict__ out)
{
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);
    plain_tile<P, SERIES>(rec, cpos, grid, kanchor, g.npx, g.ns, n0, n1, tile * P * kBlock,
                          out + static_cast<size_t>(split) * g.pstride);
}

// ---------------------------------------------------------------------------
// Plain form with the per-term Doppler window and the per-chunk skip test.
// ---------------------------------------------------------------------------

// Skip test for one pixel and chunk (reference tdcprocessor.cpp:347-383):
// probes rows max(i*64-1,0) and samples max(j*(ns/2)-1,0), i,j in {0,1,2};
// beyond the valid rows of a partial last chunk the reference's staging
// buffer still holds the previous chunk's rows, and a partial first chunk
// reads freshly allocated (zero) page-locked memory. NaN in the first probe
// makes the minimum NaN (std::min_element semantics) and skips the chunk.
__device__ __forceinline__ bool chunk_processed(const DopRec *__restrict__ rec, const Geom &g,
                                                const DopParams &dp, int chunk, float px, float py,
                                                float pz)
{
    const int valid = min(kChunk, g.npri - chunk * kChunk);
    const int sstep = g.ns / 2;
    float mn = 0.0F;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int row = max(i * (kChunk / 2) - 1, 0);
        int pri = -1;
        if (row < valid) pri = chunk * kChunk + row;
        else if (chunk > 0) pri = (chunk - 1) * kChunk + row;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int smp = max(j * sstep - 1, 0);
            float x = 0.0F, y = 0.0F, z = 0.0F, vx = 0.0F, vy = 0.0F, vz = 0.0F;
            float fc = dp.centroid ? 0.0F : dp.fc_fixed;
            if (pri >= 0) {
                const DopRec t = rec[static_cast<size_t>(pri) * g.ns + smp];
                x = t.x;
                y = t.y;
                z = t.z;
                vx = t.vx;
                vy = t.vy;
                vz = t.vz;
                fc = t.fc;
            }
            const float dx = px - x, dy = py - y, dz = pz - z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const float fd = ((vx * dx) + (vy * dy) + (vz * dz)) / __builtin_amdgcn_sqrtf(r2);
            const float mag = fabsf(fd - fc);
            if (i == 0 && j == 0) mn = mag;
            else if (mag < mn) mn = mag;
        }
    }
    return mn <= dp.half_bw;
}

template <int P, bool SERIES>
__global__ __launch_bounds__(kBlock) void doppler_kernel(const DopRec *__restrict__ rec,
                                                         const float *__restrict__ grid,
                                                         const double *__restrict__ kanchor,
                                                         Geom g, DopParams dp,
                                                         float2 *__restrict__ out,
                                                         unsigned long long *__restrict__ counters)
{
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);

    float px[P], py[P], pz[P], ar[P], ai[P];
    int pix[P];
    bool live[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        pix[j] = (tile * P + j) * kBlock + static_cast<int>(threadIdx.x);
        live[j] = pix[j] < g.npx;
        const size_t q = static_cast<size_t>(min(pix[j], g.npx - 1));
        px[j] = grid[3 * q];
        py[j] = grid[3 * q + 1];
        pz[j] = grid[3 * q + 2];
        ar[j] = 0.0F;
        ai[j] = 0.0F;
    }
    const cptr<double> ka = as_const(kanchor);
    unsigned long long processed = 0, skipped = 0;

    int n = n0;
    while (n < n1) {
        const int chunk = n / kChunk;
        const int nend = min(n1, (chunk + 1) * kChunk);
        bool on[P];
        bool any = false;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            on[j] = chunk_processed(rec, g, dp, chunk, px[j], py[j], pz[j]);
            if (live[j] && n == chunk * kChunk) {
                // count each (pixel, chunk) once: in the split holding its first chirp
                if (on[j]) ++processed;
                else ++skipped;
            }
            any |= on[j];
        }
        if (__ballot(any) != 0ULL) {
            for (int c = n; c < nend; ++c) {
                const DopRec *__restrict__ r = rec + static_cast<size_t>(c) * g.ns;
                // the phase's FP64 range from the chirp's first position (plain_tile)
                const float cx = r[0].x, cy = r[0].y, cz = r[0].z;
                double rd[P];
                float r32[P], hr[P], d0x[P], d0y[P], d0z[P];
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    rd[j] = range64(px[j], py[j], pz[j], cx, cy, cz);
                    r32[j] = static_cast<float>(rd[j]);
                    hr[j] = static_cast<float>(0.5 / rd[j]);
                    d0x[j] = px[j] - cx;
                    d0y[j] = py[j] - cy;
                    d0z[j] = pz[j] - cz;
                }
                for (int b0 = 0; b0 < g.ns; b0 += kAnchorBlock) {
                    const double kb = ka[b0 / kAnchorBlock];
                    const float kb32 = static_cast<float>(kb);
                    float thb[P];
#pragma unroll
                    for (int j = 0; j < P; ++j) thb[j] = anchor_phase(kb, rd[j]);
                    const int b1 = min(g.ns, b0 + kAnchorBlock);
#pragma unroll 2
                    for (int m = b0; m < b1; ++m) {
                        const DopRec t = r[m];
                        // offset of this sample's position from the chirp's first
                        const float ex = t.x - cx, ey = t.y - cy, ez = t.z - cz;
#pragma unroll
                        for (int j = 0; j < P; ++j) {
                            const float dx = px[j] - t.x;
                            const float dy = py[j] - t.y;
                            const float dz = pz[j] - t.z;
                            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                            const float rs = __builtin_amdgcn_rsqf(r2);
                            // windowing.h:29-37: (2/lambda) v . (p - pos) / |p - pos|
                            const float fd = ((t.vx * dx) + (t.vy * dy) + (t.vz * dz)) * rs;
                            const float delta = fd - t.fc;
                            // windowing.h:43-58; a NaN difference fails the test
                            const bool inband = on[j] && (fabsf(delta) <= dp.half_bw);
                            const float wa = dp.taper
                                                 ? fmaf(1.0F - dp.az_a, cos_rev(delta * dp.inv_bw),
                                                        dp.az_a)
                                                 : 1.0F;
                            // |d|^2 - R0^2 = -e . (d + d0), with d = d0 - e
                            const float num = -fmaf(ez, dz + d0z[j],
                                                    fmaf(ey, dy + d0y[j], ex * (dx + d0x[j])));
                            // the offset of the range (plain_tile; the expansion with SERIES)
                            float dr;
                            if constexpr (SERIES) {
                                const float x = num * hr[j];
                                dr = fmaf(-(x * hr[j]), x, x);
                            } else {
                                dr = num * __builtin_amdgcn_rcpf(r2 * rs + r32[j]);
                            }
                            const float th = __builtin_amdgcn_fractf(
                                fmaf(t.k, r32[j] + dr, fmaf(kb32, dr, thb[j])));
                            const float cs = inband ? cos_rev(th) * wa : 0.0F;
                            const float sn = inband ? sin_rev(th) * wa : 0.0F;
                            ar[j] = fmaf(t.rr, cs, ar[j]);
                            ar[j] = fmaf(-t.ri, sn, ar[j]);
                            ai[j] = fmaf(t.rr, sn, ai[j]);
                            ai[j] = fmaf(t.ri, cs, ai[j]);
                        }
                    }
                }
            }
        }
        n = nend;
    }

    float2 *o = out + static_cast<size_t>(split) * g.pstride;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        if (live[j]) o[pix[j]] = make_float2(ar[j], ai[j]);
    }
    if (processed) atomicAdd(&counters[0], processed);
    if (skipped) atomicAdd(&counters[1], skipped);
}

// Plain form on stop-and-hop input (position constant within every chirp):
// range_tile, bit-identical to plain_kernel on the same input.
template <int P>
__global__ __launch_bounds__(kBlock) void range_kernel(const float4 *__restrict__ rec,
                                                       const float4 *__restrict__ cpos,
                                                       const float *__restrict__ grid,
                                                       const double *__restrict__ kanchor, Geom g,
                                                       float2 *__restrict__ out)
{
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);
    range_tile<P, float, float4>(rec, cpos, grid, kanchor, g.npx, g.ns, n0, n1, tile * P * kBlock,
                                 out + static_cast<size_t>(split) * g.pstride);
}

// chunk_processed on stop-and-hop input: the three probe samples of a row
// hold the same position, velocity and centre, so one evaluation per row
// gives the same minimum (a NaN in the first row still makes it NaN).
__device__ __forceinline__ bool chunk_processed_chirp(const float4 *__restrict__ cpos,
                                                      const float4 *__restrict__ cvel,
                                                      const Geom &g, const DopParams &dp,
                                                      int chunk, float px, float py, float pz)
{
    const int valid = min(kChunk, g.npri - chunk * kChunk);
    float mn = 0.0F;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int row = max(i * (kChunk / 2) - 1, 0);
        int pri = -1;
        if (row < valid) pri = chunk * kChunk + row;
        else if (chunk > 0) pri = (chunk - 1) * kChunk + row;
        float4 c = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
        float4 v = make_float4(0.0F, 0.0F, 0.0F, dp.centroid ? 0.0F : dp.fc_fixed);
        if (pri >= 0) {
            c = cpos[pri];
            v = cvel[pri];
        }
        const float dx = px - c.x, dy = py - c.y, dz = pz - c.z;
        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const float fd = ((v.x * dx) + (v.y * dy) + (v.z * dz)) / __builtin_amdgcn_sqrtf(r2);
        const float mag = fabsf(fd - v.w);
        if (i == 0) mn = mag;
        else if (mag < mn) mn = mag;
    }
    return mn <= dp.half_bw;
}

// Doppler window on stop-and-hop input (position, velocity and attitude
// constant within every chirp, checked on the host): every term of a chirp
// then has the same window value for a pixel, so the window is evaluated once
// per (pixel, chirp) and scales the chirp's sum,
// sum_m w raw exp(j theta) = w sum_m raw exp(j theta), with the range of
// range_tile. A chirp that no pixel of the workgroup has in band is skipped
// (neither staged nor summed); a wave with none in band only stages.
template <int P>
__global__ __launch_bounds__(kBlock) void doppler_chirp_kernel(
    const float4 *__restrict__ rec, const float4 *__restrict__ cpos,
    const float4 *__restrict__ cvel, const float *__restrict__ grid,
    const double *__restrict__ kanchor, Geom g, DopParams dp, float2 *__restrict__ out,
    unsigned long long *__restrict__ counters)
{
    __shared__ float4 s_term[kTermStage];
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);

    float px[P], py[P], pz[P], ar[P], ai[P];
    int pix[P];
    bool live[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        pix[j] = (tile * P + j) * kBlock + static_cast<int>(threadIdx.x);
        live[j] = pix[j] < g.npx;
        const size_t q = static_cast<size_t>(min(pix[j], g.npx - 1));
        px[j] = grid[3 * q];
        py[j] = grid[3 * q + 1];
        pz[j] = grid[3 * q + 2];
        ar[j] = 0.0F;
        ai[j] = 0.0F;
    }
    unsigned long long processed = 0, skipped = 0;

    int n = n0;
    while (n < n1) {
        const int chunk = n / kChunk;
        const int nend = min(n1, (chunk + 1) * kChunk);
        bool on[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            on[j] = chunk_processed_chirp(cpos, cvel, g, dp, chunk, px[j], py[j], pz[j]);
            if (live[j] && n == chunk * kChunk) {
                if (on[j]) ++processed;
                else ++skipped;
            }
        }
        for (int c = n; c < nend; ++c) {
            const float4 cp = as_const(cpos)[c];
            const float4 cv = as_const(cvel)[c];
            double rd[P];
            float r32[P], fac[P];
            bool any = false;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                // the window of windowing.h:29-58 for this chirp (see doppler_kernel)
                const float dx = px[j] - cp.x;
                const float dy = py[j] - cp.y;
                const float dz = pz[j] - cp.z;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float rs = __builtin_amdgcn_rsqf(r2);
                const float fd = ((cv.x * dx) + (cv.y * dy) + (cv.z * dz)) * rs;
                const float delta = fd - cv.w;
                const bool inband = on[j] && (fabsf(delta) <= dp.half_bw);
                const float wa =
                    dp.taper ? fmaf(1.0F - dp.az_a, cos_rev(delta * dp.inv_bw), dp.az_a) : 1.0F;
                fac[j] = inband ? wa : 0.0F;
                // the phase's FP64 range (range_tile)
                rd[j] = range64(px[j], py[j], pz[j], cp.x, cp.y, cp.z);
                r32[j] = static_cast<float>(rd[j]);
                any |= inband;
            }
            const bool run = __ballot(any) != 0ULL;
            if (!__syncthreads_or(run ? 1 : 0)) continue;
            float tr[P], ti[P];
#pragma unroll
            for (int j = 0; j < P; ++j) {
                tr[j] = 0.0F;
                ti[j] = 0.0F;
            }
            chirp_terms<P>(rec + static_cast<size_t>(c) * g.ns, g.ns, s_term, kanchor, rd, r32, tr,
                           ti, run);
            if (!run) continue;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                ar[j] = fmaf(fac[j], tr[j], ar[j]);
                ai[j] = fmaf(fac[j], ti[j], ai[j]);
            }
        }
        n = nend;
    }

    float2 *o = out + static_cast<size_t>(split) * g.pstride;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        if (live[j]) o[pix[j]] = make_float2(ar[j], ai[j]);
    }
    if (processed) atomicAdd(&counters[0], processed);
    if (skipped) atomicAdd(&counters[1], skipped);
}

// ---------------------------------------------------------------------------
// Range-profile engine: the per-chirp tables and the interpolation kernel.
// ---------------------------------------------------------------------------

// Sample i of chirp n for the per-chirp transform: from the staged raw in the
// storage format, or (FOLD) from the complex64 input with the range window
// folded in and the storage format's rounding applied, the same values.
template <class RAW, bool FOLD>
__device__ __forceinline__ void profile_sample(const void *src, const float *wr, int n, int ns,
                                               int i, float &re, float &im)
{
    const size_t e = static_cast<size_t>(n) * ns + i;
    if constexpr (FOLD) {
        const float2 v = static_cast<const float2 *>(src)[e];
        re = v.x * wr[i];
        im = v.y * wr[i];
        RAW::round(re, im);
    } else {
        RAW::get(as_const(static_cast<const typename RAW::T *>(src) + static_cast<size_t>(n) * ns),
                 i, re, im);
    }
}

// The same sample in two steps, so that a thread issues the loads of several
// samples before it uses any: the loaded words, then their values.
template <class RAW, bool FOLD> struct ProfileWord {
    float2 v;
    float w;
};
template <class RAW> struct ProfileWord<RAW, false> {
    typename RAW::T t;
};
template <class RAW, bool FOLD>
__device__ __forceinline__ ProfileWord<RAW, FOLD> profile_load(const void *src, const float *wr,
                                                               int n, int ns, int i)
{
    const size_t e = static_cast<size_t>(n) * ns + i;
    ProfileWord<RAW, FOLD> r;
    if constexpr (FOLD) {
        r.v = static_cast<const float2 *>(src)[e];
        r.w = wr[i];
    } else {
        r.t = static_cast<const typename RAW::T *>(src)[e];
    }
    return r;
}
template <class RAW, bool FOLD>
__device__ __forceinline__ void profile_value(const ProfileWord<RAW, FOLD> &r, float &re, float &im)
{
    if constexpr (FOLD) {
        re = r.v.x * r.w;
        im = r.v.y * r.w;
        RAW::round(re, im);
    } else {
        RAW::unpack(r.t, re, im);
    }
}

// exp(+j 2 pi k / M) for k < 3 M / 4 from the half-turn table (k < M / 2):
// the second half turn is the first negated.
__device__ __forceinline__ float2 twiddle_at(const float2 *tw, int k, int half)
{
    const float2 w = tw[k < half ? k : k - half];
    return k < half ? w : make_float2(-w.x, -w.y);
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 w)
{
    return make_float2(fmaf(a.x, w.x, -(a.y * w.y)), fmaf(a.x, w.y, a.y * w.x));
}

// o[m] = sum_r a[r] w^(r m), w = exp(+j 2 pi / 8): two 4-point transforms
// of the even and odd inputs, joined by w^m
__device__ __forceinline__ void dft4_pos(float2 b0, float2 b1, float2 b2, float2 b3, float2 *o)
{
    const float2 t0 = make_float2(b0.x + b2.x, b0.y + b2.y);
    const float2 t1 = make_float2(b0.x - b2.x, b0.y - b2.y);
    const float2 t2 = make_float2(b1.x + b3.x, b1.y + b3.y);
    const float2 t3 = make_float2(b3.y - b1.y, b1.x - b3.x); // +j (b1 - b3)
    o[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
    o[1] = make_float2(t1.x + t3.x, t1.y + t3.y);
    o[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
    o[3] = make_float2(t1.x - t3.x, t1.y - t3.y);
}
__device__ __forceinline__ void dft8_pos(const float2 (&a)[8], float2 (&o)[8])
{
    constexpr float kS = 0.70710678118654752F; // 1 / sqrt(2)
    float2 ev[4], od[4];
    dft4_pos(a[0], a[2], a[4], a[6], ev);
    dft4_pos(a[1], a[3], a[5], a[7], od);
    // w^m od[m]: w = (1 + j) / sqrt 2, w^2 = j, w^3 = (-1 + j) / sqrt 2
    const float2 w1 = make_float2(kS * (od[1].x - od[1].y), kS * (od[1].x + od[1].y));
    const float2 w2 = make_float2(-od[2].y, od[2].x);
    const float2 w3 = make_float2(-kS * (od[3].x + od[3].y), kS * (od[3].x - od[3].y));
    o[0] = make_float2(ev[0].x + od[0].x, ev[0].y + od[0].y);
    o[4] = make_float2(ev[0].x - od[0].x, ev[0].y - od[0].y);
    o[1] = make_float2(ev[1].x + w1.x, ev[1].y + w1.y);
    o[5] = make_float2(ev[1].x - w1.x, ev[1].y - w1.y);
    o[2] = make_float2(ev[2].x + w2.x, ev[2].y + w2.y);
    o[6] = make_float2(ev[2].x - w2.x, ev[2].y - w2.y);
    o[3] = make_float2(ev[3].x + w3.x, ev[3].y + w3.y);
    o[7] = make_float2(ev[3].x - w3.x, ev[3].y - w3.y);
}

// Row n of the table from the transform x in LDS: row[j] = x[(j - lead) mod M]
// (x[(lead - j) mod M] for down-chirps) for j < kProfileDeltaEntry past M,
// then the chirp's position increment against chirp n - 1 (zero for the
// first chirp) as one float4 in the last two entries, rounded from FP64:
// (c_n - c_{n-1}, (|c_n|^2 - |c_{n-1}|^2) / 2).
template <int BLOCK>
__device__ __forceinline__ void write_profile_row(const float2 *x, const ChirpPos *__restrict__ cpos,
                                                  int n, int M, int Mp, int lead, int reverse,
                                                  float2 *__restrict__ table)
{
    const int mask = M - 1;
    const int tid = static_cast<int>(threadIdx.x);
    float2 *dst = table + static_cast<size_t>(n) * Mp;
    const int rows = M + kProfileDeltaEntry;
    for (int j = tid; j < rows; j += BLOCK) dst[j] = x[(reverse ? lead - j : j - lead) & mask];
    if (tid == 0) {
        float4 d = make_float4(0.0F, 0.0F, 0.0F, 0.0F);
        if (n > 0) {
            const ChirpPos c = cpos[n], p = cpos[n - 1];
            d = make_float4(static_cast<float>(c.x - p.x), static_cast<float>(c.y - p.y),
                            static_cast<float>(c.z - p.z), static_cast<float>(0.5 * (c.pad - p.pad)));
        }
        *reinterpret_cast<float4 *>(dst + rows) = d;
    }
}

// One workgroup per chirp: b[(i - h) mod M] = c_i * corr[i] (zero elsewhere),
// then B[m] = sum_j b[j] exp(+j 2 pi j m / M) by radix-8 Stockham passes (and
// one last radix-4 or radix-2 pass) between two LDS buffers, with the
// twiddles staged in LDS up to kProfileLdsTwiddleM (the output comes out in
// natural order), written
// shifted by lead = taps / 2 - 1 bins, row[j] = B[(j - lead) mod M], over M +
// kProfileDeltaEntry entries: a pixel whose bin position has integer part f
// reads its taps B[f - lead .. f - lead + taps - 1] as row[f .. f + taps - 1],
// and a wave's 64-entry window starting at bin l as row[l .. l + 63], which
// never wrap. For down-chirps (`reverse`) the row holds
// B[(lead - j) mod M]: the profile at frequency -u is read at u. The row's
// last two entries hold the chirp's position increment (write_profile_row).
template <class RAW, int BLOCK, bool FOLD, bool LDS_TW>
__global__ __launch_bounds__(BLOCK) void stage_profile_kernel(
    const void *__restrict__ src, const float *__restrict__ wr, const float *__restrict__ corr,
    const float2 *__restrict__ twiddle, const ChirpPos *__restrict__ cpos, int ns, int M, int Mp,
    int lead, int reverse, float2 *__restrict__ table)
{
    // 2 M entries (+ M / 2 twiddles with LDS_TW), sized at launch
    extern __shared__ float2 buf[];
    const int n = static_cast<int>(blockIdx.x);
    const int tid = static_cast<int>(threadIdx.x);
    const int h = ns / 2;
    const int half = M / 2;
    float2 *x = buf;
    float2 *y = buf + M;
    const float2 *tw = twiddle;
    const int mask = M - 1;
    // the thread's first kTwiddles twiddles, loaded beside its first samples
    constexpr int kTwiddles = 4;
    float2 twr[kTwiddles];
    if constexpr (LDS_TW) {
#pragma unroll
        for (int u = 0; u < kTwiddles; ++u) twr[u] = twiddle[min(tid + u * BLOCK, half - 1)];
    }
    // b[j] = c_i corr[i] for j = (i - h) mod M: zeros at j in [ns - h, M - h),
    // then the ns samples, kUnroll per thread at a time with their loads
    // issued before the first is used
    for (int j = ns - h + tid; j < M - h; j += BLOCK) x[j] = make_float2(0.0F, 0.0F);
    constexpr int kUnroll = 4;
    auto samples = [&](int i0) {
        ProfileWord<RAW, FOLD> wd[kUnroll];
        float c[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int i = min(i0 + u * BLOCK, ns - 1);
            wd[u] = profile_load<RAW, FOLD>(src, wr, n, ns, i);
            c[u] = corr[i];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const int i = i0 + u * BLOCK;
            if (i < ns) {
                float re, im;
                profile_value<RAW, FOLD>(wd[u], re, im);
                x[(i - h) & mask] = make_float2(re * c[u], im * c[u]);
            }
        }
    };
    // (the first batch outside the loop, so that its loads issue with the
    // twiddles')
    if (ns > 0) samples(tid);
    for (int i0 = tid + kUnroll * BLOCK; i0 < ns; i0 += kUnroll * BLOCK) samples(i0);
    if constexpr (LDS_TW) {
        float2 *t = buf + 2 * M;
#pragma unroll
        for (int u = 0; u < kTwiddles; ++u) {
            if (tid + u * BLOCK < half) t[tid + u * BLOCK] = twr[u];
        }
        for (int k = tid + kTwiddles * BLOCK; k < half; k += BLOCK) t[k] = twiddle[k];
        tw = t;
    }
    __syncthreads();
    // Stockham passes, B[m] = sum_j b[j] exp(+j 2 pi j m / M) in natural order:
    // radix 8 while an eighth of the remaining length L = M / s is left, then
    // one radix-4 or radix-2 pass for the last factor of 4 or 2
    int s = 1, lg = 0;
    for (; (M >> lg) >= 8; s <<= 3, lg += 3) {
        const int q = (M >> lg) >> 3; // L / 8
        for (int idx = tid; idx < M / 8; idx += BLOCK) {
            const int j = idx >> lg;
            const int k = idx & (s - 1);
            float2 a[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) a[r] = x[k + s * (j + r * q)];
            float2 o[8];
            dft8_pos(a, o);
            const int e = j * s;
            float2 *dst = y + k + 8 * s * j;
            dst[0] = o[0];
#pragma unroll
            for (int m = 1; m < 8; ++m) dst[m * s] = cmulf(o[m], twiddle_at(tw, m * e, half));
        }
        __syncthreads();
        float2 *tmp = x;
        x = y;
        y = tmp;
    }
    if ((M >> lg) == 4) {
        // the last radix-4 pass: L = 4, so every twiddle is 1
        for (int k = tid; k < M / 4; k += BLOCK) {
            const float2 a0 = x[k], a1 = x[k + s], a2 = x[k + 2 * s], a3 = x[k + 3 * s];
            const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y);
            const float2 t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
            const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y);
            // +j (a1 - a3): the quarter turn of the +j 2 pi transform
            const float2 t3 = make_float2(a3.y - a1.y, a1.x - a3.x);
            y[k] = make_float2(t0.x + t2.x, t0.y + t2.y);
            y[k + s] = make_float2(t1.x + t3.x, t1.y + t3.y);
            y[k + 2 * s] = make_float2(t0.x - t2.x, t0.y - t2.y);
            y[k + 3 * s] = make_float2(t1.x - t3.x, t1.y - t3.y);
        }
        __syncthreads();
        float2 *tmp = x;
        x = y;
        y = tmp;
    } else if ((M >> lg) == 2) {
        for (int k = tid; k < half; k += BLOCK) {
            const float2 a = x[k], c = x[k + half];
            y[k] = make_float2(a.x + c.x, a.y + c.y);
            y[k + half] = make_float2(a.x - c.x, a.y - c.y);
        }
        __syncthreads();
        float2 *tmp = x;
        x = y;
        y = tmp;
    }
    write_profile_row<BLOCK>(x, cpos, n, M, Mp, lead, reverse, table);
}

// The same transform for chirps whose two Stockham buffers would not fit the
// LDS (M above kProfileStockhamMaxM): in place in one buffer, radix-2
// decimation in time from a bit-reversed placement of b, then written like
// the Stockham kernel's rows.
template <class RAW, int BLOCK, bool FOLD>
__global__ __launch_bounds__(BLOCK) void stage_profile_inplace_kernel(
    const void *__restrict__ src, const float *__restrict__ wr, const float *__restrict__ corr,
    const float2 *__restrict__ twiddle, const ChirpPos *__restrict__ cpos, int ns, int M, int Mp,
    int lead, int reverse, float2 *__restrict__ table)
{
    extern __shared__ float2 x[]; // M entries, sized at launch
    const int n = static_cast<int>(blockIdx.x);
    const int tid = static_cast<int>(threadIdx.x);
    const int h = ns / 2;
    const int mask = M - 1;
    const int bits = __builtin_ctz(static_cast<unsigned>(M));
    for (int j = tid; j < M; j += BLOCK) x[j] = make_float2(0.0F, 0.0F);
    __syncthreads();
    for (int i = tid; i < ns; i += BLOCK) {
        float re, im;
        profile_sample<RAW, FOLD>(src, wr, n, ns, i, re, im);
        const float c = corr[i];
        const unsigned j = static_cast<unsigned>((i - h) & mask);
        x[__builtin_bitreverse32(j) >> (32 - bits)] = make_float2(re * c, im * c);
    }
    __syncthreads();
    for (int lg = 0; (2 << lg) <= M; ++lg) {
        // butterflies of span 2^(lg+1): (base, base + half), twiddle exp(+j 2 pi k / M)
        const int half = 1 << lg;
        const int stride = M >> (lg + 1);
        for (int idx = tid; idx < M / 2; idx += BLOCK) {
            const int k = idx & (half - 1);
            const int base = ((idx >> lg) << (lg + 1)) + k;
            const float2 w = twiddle[k * stride];
            const float2 a = x[base];
            const float2 d = x[base + half];
            const float cr = fmaf(d.x, w.x, -(d.y * w.y));
            const float ci = fmaf(d.x, w.y, d.y * w.x);
            x[base] = make_float2(a.x + cr, a.y + ci);
            x[base + half] = make_float2(a.x - cr, a.y - ci);
        }
        __syncthreads();
    }
    write_profile_row<BLOCK>(x, cpos, n, M, Mp, lead, reverse, table);
}

template <int TAPS, int OVER>
__global__ __launch_bounds__(kBlock) void profile_kernel(const float2 *__restrict__ table,
                                                         const ChirpPos *__restrict__ cpos,
                                                         const double *__restrict__ grid,
                                                         ProfileParams pp, Geom g,
                                                         float2 *__restrict__ out)
{
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);
    profile_tile<TAPS, OVER>(table, cpos, grid, pp, g.npx, n0, n1, tile * kBlock,
                             out + static_cast<size_t>(split) * g.pstride);
}

// Sum of the per-split partial images in split order.
__global__ __launch_bounds__(kBlock) void finalize_f32_kernel(const float2 *__restrict__ partial,
                                                              int nsplit, int npx,
                                                              float2 *__restrict__ img)
{
    const int p = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    if (p >= npx) return;
    float sr = 0.0F, si = 0.0F;
    for (int s = 0; s < nsplit; ++s) {
        const float2 v = partial[static_cast<size_t>(s) * npx + p];
        sr += v.x;
        si += v.y;
    }
    img[p] = make_float2(sr, si);
}

// ---------------------------------------------------------------------------
// Precision form.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void stage_precision_kernel(const float2 *__restrict__ raw,
                                                                 const float *__restrict__ wr,
                                                                 int npri, int ns, int storage,
                                                                 void *__restrict__ out)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    for (size_t e = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * kBlock) {
        const int m = static_cast<int>(e % ns);
        const float2 v = raw[e];
        const float re = v.x * wr[m], im = v.y * wr[m];
        if (storage == 0) {
            static_cast<float2 *>(out)[e] = make_float2(re, im);
        } else if (storage == 1) {
            const unsigned short hr = __builtin_bit_cast(unsigned short, static_cast<_Float16>(re));
            const unsigned short hi = __builtin_bit_cast(unsigned short, static_cast<_Float16>(im));
            static_cast<unsigned int *>(out)[e] = static_cast<unsigned int>(hr) |
                                                  (static_cast<unsigned int>(hi) << 16);
        } else {
            // round to nearest even (finite inputs)
            unsigned int br = __float_as_uint(re), bi = __float_as_uint(im);
            br += 0x7fffU + ((br >> 16) & 1U);
            bi += 0x7fffU + ((bi >> 16) & 1U);
            static_cast<unsigned int *>(out)[e] = (br >> 16) | (bi & 0xffff0000U);
        }
    }
}

template <int P, class RAW>
__global__ __launch_bounds__(kBlock) void precision_kernel(
    const typename RAW::T *__restrict__ raw, const ChirpPos *__restrict__ cpos,
    const double *__restrict__ grid, const double *__restrict__ kanchor, double dk, Geom g,
    double2 *__restrict__ out)
{
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);
    precision_tile<P, RAW>(raw, cpos, grid, kanchor, dk, g.npx, g.ns, n0, n1, tile * P * kBlock,
                           out + static_cast<size_t>(split) * g.pstride);
}

// Matrix engine staging: one workgroup per chirp. The chirp's samples (in
// their storage format) are scaled by a power of two that puts the largest
// component in [2^13, 2^14) and laid out as the A operand of
// v_mfma_f32_32x32x16_f16 per 256-sample tile set q and k-slice s: lane L
// holds row m = L % 32 and contraction half hk = L / 32; row m is (block
// 2q + (m/4)%2 of 128 samples, group a = 2 (m/8) + (m%4)/2 of 16 samples,
// real part for even m); element j of the lane pairs with w^c, c = 8s + j,
// Re w^c for hk = 0 and Im w^c for hk = 1, so a real row holds (Re x, -Im x)
// and an imaginary row (Im x, Re x) for x = raw[128 block + 16a + c].
template <class RAW>
__device__ __forceinline__ void load_raw(const typename RAW::T *p, size_t e, float &re, float &im);
template <>
__device__ __forceinline__ void load_raw<RawC64>(const float2 *p, size_t e, float &re, float &im)
{
    const float2 v = p[e];
    re = v.x;
    im = v.y;
}
template <>
__device__ __forceinline__ void load_raw<RawF16>(const unsigned int *p, size_t e, float &re,
                                                 float &im)
{
    const unsigned int v = p[e];
    re = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(v & 0xffffU)));
    im = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(v >> 16)));
}
template <>
__device__ __forceinline__ void load_raw<RawBF16>(const unsigned int *p, size_t e, float &re,
                                                  float &im)
{
    const unsigned int v = p[e];
    re = __uint_as_float(v << 16);
    im = __uint_as_float(v & 0xffff0000U);
}

template <class RAW, int PARTS>
__global__ __launch_bounds__(kBlock) void stage_matrix_kernel(const typename RAW::T *__restrict__ raw,
                                                              int npri, int ns,
                                                              uintx4_t *__restrict__ atab,
                                                              float *__restrict__ inv_scale)
{
    __shared__ float s_max[kBlock / 64];
    const int nq = matrix_quads(ns);
    for (int n = static_cast<int>(blockIdx.x); n < npri; n += static_cast<int>(gridDim.x)) {
        const size_t row = static_cast<size_t>(n) * ns;
        float mx = 0.0F;
        for (int m = static_cast<int>(threadIdx.x); m < ns; m += kBlock) {
            float re, im;
            load_raw<RAW>(raw, row + m, re, im);
            mx = fmaxf(mx, fmaxf(fabsf(re), fabsf(im)));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        __syncthreads();
        int sh = 0;
        if (mx > 0.0F) {
            int e;
            frexpf(mx, &e);
            sh = 14 - max(e, -100);
        }
        if (threadIdx.x == 0) inv_scale[n] = ldexpf(1.0F, -sh);
        // one item = (tile set, k-slice, lane)
        for (int it = static_cast<int>(threadIdx.x); it < nq * 2 * 64; it += kBlock) {
            const int qs = it >> 6, L = it & 63;
            const int q = qs >> 1, sl = qs & 1;
            const int m = L & 31, hk = L >> 5;
            const int blk = 2 * q + ((m >> 2) & 1);
            const int a = 2 * (m >> 3) + ((m & 3) >> 1);
            const bool imag_row = (m & 1) != 0;
            half8_t hi, lo;
#pragma unroll
            for (int j = 0; j < kSlice; ++j) {
                const int i = blk * kMatrixBlock + kGroup * a + kSlice * sl + j;
                float re = 0.0F, im = 0.0F;
                if (i < ns) load_raw<RAW>(raw, row + i, re, im);
                re = ldexpf(re, sh);
                im = ldexpf(im, sh);
                const float v = imag_row ? (hk ? re : im) : (hk ? -im : re);
                const _Float16 h = static_cast<_Float16>(v);
                hi[j] = h;
                lo[j] = static_cast<_Float16>(v - static_cast<float>(h));
            }
            uintx4_t *dst = atab + ((static_cast<size_t>(n) * nq + q) * 2 * PARTS + sl * PARTS) * 64 + L;
            dst[0] = __builtin_bit_cast(uintx4_t, hi);
            if (PARTS == 2) dst[64] = __builtin_bit_cast(uintx4_t, lo);
        }
    }
}

template <int PARTS>
__global__ __launch_bounds__(kBlock) void precision_matrix_kernel(
    const uintx4_t *__restrict__ atab, const float *__restrict__ inv_scale,
    const ChirpPos *__restrict__ cpos, const double *__restrict__ grid,
    const double *__restrict__ kanchor, double dk, Geom g, double2 *__restrict__ out)
{
    int tile, split;
    map_block(g, tile, split);
    if (tile >= g.ntiles) return;
    const int n0 = split * g.cps;
    const int n1 = min(g.npri, n0 + g.cps);
    matrix_tile<PARTS>(atab, inv_scale, cpos, grid, kanchor, dk, g.npx, matrix_quads(g.ns), n0, n1,
                       tile * kBlock, out + static_cast<size_t>(split) * g.pstride);
}

// Adds pixel p of the nsplit partial images (npx apart) to (sr, si) in split
// order, with the loads of kSplitBatch splits in flight at a time.
template <class PT>
__device__ __forceinline__ void add_splits(const PT *__restrict__ partial, int nsplit, int npx,
                                           int p, double &sr, double &si)
{
    constexpr int kSplitBatch = 8;
    for (int s0 = 0; s0 < nsplit; s0 += kSplitBatch) {
        PT v[kSplitBatch];
#pragma unroll
        for (int u = 0; u < kSplitBatch; ++u)
            v[u] = partial[static_cast<size_t>(min(s0 + u, nsplit - 1)) * npx + p];
#pragma unroll
        for (int u = 0; u < kSplitBatch; ++u) {
            if (s0 + u < nsplit) {
                sr += static_cast<double>(v[u].x);
                si += static_cast<double>(v[u].y);
            }
        }
    }
}

// FP64 sum of the per-split partial images (FP64, or FP32 from the
// range-profile engine) in split order.
template <class PT>
__global__ __launch_bounds__(kBlock) void finalize_f64_kernel(const PT *__restrict__ partial,
                                                              int nsplit, int npx,
                                                              float2 *__restrict__ img)
{
    const int p = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    if (p >= npx) return;
    double sr = 0.0, si = 0.0;
    add_splits(partial, nsplit, npx, p, sr, si);
    img[p] = make_float2(static_cast<float>(sr), static_cast<float>(si));
}

int stage_grid(size_t total)
{
    const size_t want = (total + kBlock - 1) / kBlock;
    return static_cast<int>(want < 4096 ? (want > 0 ? want : 1) : 4096);
}

int correlate_grid(const Geom &g) { return g.tiles_padded * g.nsplit; }

} // namespace

hipError_t launch_stage_plain(const float *raw, const float *pos, const float *dk, const float *wr,
                              int npri, int ns, PlainRec *out, hipStream_t s)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    hipLaunchKernelGGL(stage_plain_kernel, dim3(stage_grid(total)), dim3(kBlock), 0, s,
                       reinterpret_cast<const float2 *>(raw), pos, dk, wr, npri, ns, out);
    return hipGetLastError();
}

hipError_t launch_stage_doppler(const float *raw, const float *pos, const float *vel,
                                const float *att, const float *dk, const float *wr, int npri,
                                int ns, float two_over_lambda, const float *bore, int centroid,
                                float fc_fixed, DopRec *out, hipStream_t s)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    hipLaunchKernelGGL(stage_doppler_kernel, dim3(stage_grid(total)), dim3(kBlock), 0, s,
                       reinterpret_cast<const float2 *>(raw), pos, vel,
                       reinterpret_cast<const float4 *>(att), dk, wr, npri, ns, two_over_lambda,
                       make_float3(bore[0], bore[1], bore[2]), centroid, fc_fixed, out);
    return hipGetLastError();
}

namespace {

template <bool SERIES>
hipError_t plain_launch(int ppl, int nb, const PlainRec *rec, const float4 *cpos,
                        const float *grid, const double *kanchor, const Geom &g, float2 *out,
                        hipStream_t s)
{
    switch (ppl) {
    case 1:
        hipLaunchKernelGGL((plain_kernel<1, SERIES>), dim3(nb), dim3(kBlock), 0, s, rec, cpos, grid,
                           kanchor, g, out);
        break;
    case 2:
        hipLaunchKernelGGL((plain_kernel<2, SERIES>), dim3(nb), dim3(kBlock), 0, s, rec, cpos, grid,
                           kanchor, g, out);
        break;
    case 4:
        hipLaunchKernelGGL((plain_kernel<4, SERIES>), dim3(nb), dim3(kBlock), 0, s, rec, cpos, grid,
                           kanchor, g, out);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool SERIES>
hipError_t doppler_launch(int ppl, int nb, const DopRec *rec, const float *grid,
                          const double *kanchor, const Geom &g, const DopParams &dp, float2 *out,
                          unsigned long long *counters, hipStream_t s)
{
    switch (ppl) {
    case 1:
        hipLaunchKernelGGL((doppler_kernel<1, SERIES>), dim3(nb), dim3(kBlock), 0, s, rec, grid,
                           kanchor, g, dp, out, counters);
        break;
    case 2:
        hipLaunchKernelGGL((doppler_kernel<2, SERIES>), dim3(nb), dim3(kBlock), 0, s, rec, grid,
                           kanchor, g, dp, out, counters);
        break;
    case 4:
        hipLaunchKernelGGL((doppler_kernel<4, SERIES>), dim3(nb), dim3(kBlock), 0, s, rec, grid,
                           kanchor, g, dp, out, counters);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace

hipError_t launch_plain(int ppl, bool series, const PlainRec *rec, const float4 *cpos,
                        const float *grid, const double *kanchor, const Geom &g, float2 *out,
                        hipStream_t s, int *workgroups)
{
    const int nb = correlate_grid(g);
    *workgroups = nb;
    return series ? plain_launch<true>(ppl, nb, rec, cpos, grid, kanchor, g, out, s)
                  : plain_launch<false>(ppl, nb, rec, cpos, grid, kanchor, g, out, s);
}

hipError_t launch_doppler(int ppl, bool series, const DopRec *rec, const float *grid,
                          const double *kanchor, const Geom &g, const DopParams &dp, float2 *out,
                          unsigned long long *counters, hipStream_t s, int *workgroups)
{
    const int nb = correlate_grid(g);
    *workgroups = nb;
    return series ? doppler_launch<true>(ppl, nb, rec, grid, kanchor, g, dp, out, counters, s)
                  : doppler_launch<false>(ppl, nb, rec, grid, kanchor, g, dp, out, counters, s);
}

hipError_t launch_stage_terms(const float *raw, const float *dk, const float *wr, int npri, int ns,
                              float4 *out, hipStream_t s)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    hipLaunchKernelGGL(stage_terms_kernel, dim3(stage_grid(total)), dim3(kBlock), 0, s,
                       reinterpret_cast<const float2 *>(raw), dk, wr, npri, ns, out);
    return hipGetLastError();
}

hipError_t launch_stage_chirps(const float *pos, const float *vel, const float *att, int npri,
                               int ns, float two_over_lambda, const float *bore, int centroid,
                               float fc_fixed, float4 *cpos, float4 *cvel, hipStream_t s)
{
    const float3 b = bore ? make_float3(bore[0], bore[1], bore[2]) : make_float3(0.0F, 0.0F, 0.0F);
    hipLaunchKernelGGL(stage_chirps_kernel, dim3(stage_grid(static_cast<size_t>(npri))),
                       dim3(kBlock), 0, s, pos, vel, reinterpret_cast<const float4 *>(att), npri,
                       ns, two_over_lambda, b, centroid, fc_fixed, cpos, cvel);
    return hipGetLastError();
}

hipError_t launch_range(int ppl, const float4 *rec, const float4 *cpos, const float *grid,
                        const double *kanchor, const Geom &g, float2 *out, hipStream_t s,
                        int *workgroups)
{
    const int nb = correlate_grid(g);
    *workgroups = nb;
    switch (ppl) {
    case 1:
        hipLaunchKernelGGL(range_kernel<1>, dim3(nb), dim3(kBlock), 0, s, rec, cpos, grid, kanchor,
                           g, out);
        break;
    case 2:
        hipLaunchKernelGGL(range_kernel<2>, dim3(nb), dim3(kBlock), 0, s, rec, cpos, grid, kanchor,
                           g, out);
        break;
    case 4:
        hipLaunchKernelGGL(range_kernel<4>, dim3(nb), dim3(kBlock), 0, s, rec, cpos, grid, kanchor,
                           g, out);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_doppler_chirp(int ppl, const float4 *rec, const float4 *cpos,
                                const float4 *cvel, const float *grid, const double *kanchor,
                                const Geom &g, const DopParams &dp, float2 *out,
                                unsigned long long *counters, hipStream_t s, int *workgroups)
{
    const int nb = correlate_grid(g);
    *workgroups = nb;
    switch (ppl) {
    case 1:
        hipLaunchKernelGGL(doppler_chirp_kernel<1>, dim3(nb), dim3(kBlock), 0, s, rec, cpos, cvel,
                           grid, kanchor, g, dp, out, counters);
        break;
    case 2:
        hipLaunchKernelGGL(doppler_chirp_kernel<2>, dim3(nb), dim3(kBlock), 0, s, rec, cpos, cvel,
                           grid, kanchor, g, dp, out, counters);
        break;
    case 4:
        hipLaunchKernelGGL(doppler_chirp_kernel<4>, dim3(nb), dim3(kBlock), 0, s, rec, cpos, cvel,
                           grid, kanchor, g, dp, out, counters);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

namespace {

// Twiddles are staged in LDS with the two Stockham buffers up to this
// transform length (2.5 M entries, 80 KiB at 4096); beyond, they are read from
// the table in global memory
constexpr int kProfileLdsTwiddleM = 4096;

template <class RAW, int BLOCK, bool FOLD>
hipError_t stage_profile_launch(const void *src, const float *wr, const float *corr,
                                const float2 *twiddle, const ChirpPos *cpos, int npri, int ns,
                                int M, int Mp, int lead, int reverse, float2 *table, hipStream_t s)
{
    const bool inplace = M > kProfileStockhamMaxM;
    const bool ldstw = M <= kProfileLdsTwiddleM;
    const size_t lds = (inplace ? 2 : (ldstw ? 5 : 4)) * sizeof(float2) * static_cast<size_t>(M) / 2;
    const void *kernel =
        inplace ? reinterpret_cast<const void *>(&stage_profile_inplace_kernel<RAW, BLOCK, FOLD>)
        : ldstw ? reinterpret_cast<const void *>(&stage_profile_kernel<RAW, BLOCK, FOLD, true>)
                : reinterpret_cast<const void *>(&stage_profile_kernel<RAW, BLOCK, FOLD, false>);
    // more than 64 KiB of dynamic LDS is granted to a kernel on request
    if (lds > (size_t(64) << 10)) {
        const hipError_t e = hipFuncSetAttribute(
            kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    if (inplace)
        hipLaunchKernelGGL((stage_profile_inplace_kernel<RAW, BLOCK, FOLD>), dim3(npri),
                           dim3(BLOCK), lds, s, src, wr, corr, twiddle, cpos, ns, M, Mp, lead,
                           reverse, table);
    else if (ldstw)
        hipLaunchKernelGGL((stage_profile_kernel<RAW, BLOCK, FOLD, true>), dim3(npri), dim3(BLOCK),
                           lds, s, src, wr, corr, twiddle, cpos, ns, M, Mp, lead, reverse, table);
    else
        hipLaunchKernelGGL((stage_profile_kernel<RAW, BLOCK, FOLD, false>), dim3(npri), dim3(BLOCK),
                           lds, s, src, wr, corr, twiddle, cpos, ns, M, Mp, lead, reverse, table);
    return hipGetLastError();
}

// One workgroup per chirp. Long transforms take most of a CU's LDS, so only
// one workgroup fits per CU: they run 1024 threads (16 waves) to keep the
// LDS busy across the passes' barriers; short ones run 256.
template <class RAW, bool FOLD>
hipError_t stage_profile_raw(const void *src, const float *wr, const float *corr,
                             const float2 *twiddle, const ChirpPos *cpos, int npri, int ns, int M,
                             int Mp, int lead, int reverse, float2 *table, hipStream_t s)
{
    if (M >= kProfileWideM)
        return stage_profile_launch<RAW, 1024, FOLD>(src, wr, corr, twiddle, cpos, npri, ns, M,
                                                     Mp, lead, reverse, table, s);
    return stage_profile_launch<RAW, kBlock, FOLD>(src, wr, corr, twiddle, cpos, npri, ns, M, Mp,
                                                   lead, reverse, table, s);
}

template <bool FOLD>
hipError_t stage_profile_any(int storage, const void *src, const float *wr, const float *corr,
                             const float2 *twiddle, const ChirpPos *cpos, int npri, int ns, int M,
                             int Mp, int taps, int reverse, float2 *table, hipStream_t s)
{
    // taps reach row[M + taps - 2]; the position increment sits after
    // kProfileDeltaEntry entries past M
    if (M > kProfileMaxM || M < 2 || (M & (M - 1)) != 0 || taps < 2 || taps > kProfileTaps ||
        Mp != M + kProfilePad || M + taps - 1 > M + kProfileDeltaEntry || !cpos)
        return hipErrorInvalidValue;
    const int lead = taps / 2 - 1;
    switch (storage) {
    case 0:
        return stage_profile_raw<RawC64, FOLD>(src, wr, corr, twiddle, cpos, npri, ns, M, Mp, lead,
                                               reverse, table, s);
    case 1:
        return stage_profile_raw<RawF16, FOLD>(src, wr, corr, twiddle, cpos, npri, ns, M, Mp, lead,
                                               reverse, table, s);
    case 2:
        return stage_profile_raw<RawBF16, FOLD>(src, wr, corr, twiddle, cpos, npri, ns, M, Mp,
                                                lead, reverse, table, s);
    default:
        return hipErrorInvalidValue;
    }
}

} // namespace

hipError_t launch_stage_profile(int storage, const void *staged, const float *corr,
                                const float2 *twiddle, const ChirpPos *cpos, int npri, int ns,
                                int M, int Mp, int taps, int reverse, float2 *table, hipStream_t s)
{
    return stage_profile_any<false>(storage, staged, nullptr, corr, twiddle, cpos, npri, ns, M, Mp,
                                    taps, reverse, table, s);
}

hipError_t launch_stage_profile_folded(int storage, const float *raw, const float *wr,
                                       const float *corr, const float2 *twiddle,
                                       const ChirpPos *cpos, int npri, int ns, int M, int Mp,
                                       int taps, int reverse, float2 *table, hipStream_t s)
{
    return stage_profile_any<true>(storage, raw, wr, corr, twiddle, cpos, npri, ns, M, Mp, taps,
                                   reverse, table, s);
}

hipError_t launch_profile(const float2 *table, const ChirpPos *cpos, const double *grid,
                          const ProfileParams &pp, const Geom &g, float2 *out, hipStream_t s,
                          int *workgroups)
{
    const int nb = correlate_grid(g);
    *workgroups = nb;
    if (pp.over == 4 && pp.taps == 6)
        hipLaunchKernelGGL((profile_kernel<6, 4>), dim3(nb), dim3(kBlock), 0, s, table, cpos, grid,
                           pp, g, out);
    else if (pp.over == 4 && pp.taps == 8)
        hipLaunchKernelGGL((profile_kernel<8, 4>), dim3(nb), dim3(kBlock), 0, s, table, cpos, grid,
                           pp, g, out);
    else if (pp.over == 2 && pp.taps == 8)
        hipLaunchKernelGGL((profile_kernel<8, 2>), dim3(nb), dim3(kBlock), 0, s, table, cpos, grid,
                           pp, g, out);
    else if (pp.over == 2 && pp.taps == 10)
        hipLaunchKernelGGL((profile_kernel<10, 2>), dim3(nb), dim3(kBlock), 0, s, table, cpos, grid,
                           pp, g, out);
    else if (pp.over == 2 && pp.taps == 12)
        hipLaunchKernelGGL((profile_kernel<12, 2>), dim3(nb), dim3(kBlock), 0, s, table, cpos, grid,
                           pp, g, out);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_finalize_f32(const float2 *partial, int nsplit, int npx, float2 *img,
                               hipStream_t s)
{
    hipLaunchKernelGGL(finalize_f32_kernel, dim3((npx + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       partial, nsplit, npx, img);
    return hipGetLastError();
}

hipError_t launch_stage_precision(int storage, const float *raw, const float *wr, int npri, int ns,
                                  void *out, hipStream_t s)
{
    const size_t total = static_cast<size_t>(npri) * ns;
    hipLaunchKernelGGL(stage_precision_kernel, dim3(stage_grid(total)), dim3(kBlock), 0, s,
                       reinterpret_cast<const float2 *>(raw), wr, npri, ns, storage, out);
    return hipGetLastError();
}

template <class RAW>
static hipError_t launch_precision_t(int ppl, const void *raw, const ChirpPos *cpos,
                                     const double *grid, const double *kanchor, double dk,
                                     const Geom &g, double2 *out, hipStream_t s)
{
    const int nb = correlate_grid(g);
    const auto *r = static_cast<const typename RAW::T *>(raw);
    switch (ppl) {
    case 1:
        hipLaunchKernelGGL((precision_kernel<1, RAW>), dim3(nb), dim3(kBlock), 0, s, r, cpos, grid,
                           kanchor, dk, g, out);
        break;
    case 2:
        hipLaunchKernelGGL((precision_kernel<2, RAW>), dim3(nb), dim3(kBlock), 0, s, r, cpos, grid,
                           kanchor, dk, g, out);
        break;
    case 4:
        hipLaunchKernelGGL((precision_kernel<4, RAW>), dim3(nb), dim3(kBlock), 0, s, r, cpos, grid,
                           kanchor, dk, g, out);
        break;
    case 8:
        hipLaunchKernelGGL((precision_kernel<8, RAW>), dim3(nb), dim3(kBlock), 0, s, r, cpos, grid,
                           kanchor, dk, g, out);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_precision(int ppl, int storage, const void *raw, const ChirpPos *cpos,
                            const double *grid, const double *kanchor, double dk, const Geom &g,
                            double2 *out, hipStream_t s, int *workgroups)
{
    *workgroups = correlate_grid(g);
    switch (storage) {
    case 0:
        return launch_precision_t<RawC64>(ppl, raw, cpos, grid, kanchor, dk, g, out, s);
    case 1:
        return launch_precision_t<RawF16>(ppl, raw, cpos, grid, kanchor, dk, g, out, s);
    case 2:
        return launch_precision_t<RawBF16>(ppl, raw, cpos, grid, kanchor, dk, g, out, s);
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_stage_matrix(int storage, const void *staged, int npri, int ns, int parts,
                               void *atab, float *inv_scale, hipStream_t s)
{
    if (parts != matrix_parts(storage)) return hipErrorInvalidValue;
    const int grid = npri < 65535 ? (npri > 0 ? npri : 1) : 65535;
    auto *t = static_cast<uintx4_t *>(atab);
    switch (storage) {
    case 0:
        hipLaunchKernelGGL((stage_matrix_kernel<RawC64, 2>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const float2 *>(staged), npri, ns, t, inv_scale);
        break;
    case 1:
        hipLaunchKernelGGL((stage_matrix_kernel<RawF16, 1>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const unsigned int *>(staged), npri, ns, t, inv_scale);
        break;
    case 2:
        hipLaunchKernelGGL((stage_matrix_kernel<RawBF16, 1>), dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const unsigned int *>(staged), npri, ns, t, inv_scale);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_precision_matrix(int parts, const void *atab, const float *inv_scale,
                                   const ChirpPos *cpos, const double *grid,
                                   const double *kanchor, double dk, const Geom &g,
                                   double2 *out, hipStream_t s, int *workgroups)
{
    const int nb = correlate_grid(g);
    *workgroups = nb;
    const auto *t = static_cast<const uintx4_t *>(atab);
    if (parts == 2) {
        hipLaunchKernelGGL(precision_matrix_kernel<2>, dim3(nb), dim3(kBlock), 0, s, t, inv_scale,
                           cpos, grid, kanchor, dk, g, out);
    } else if (parts == 1) {
        hipLaunchKernelGGL(precision_matrix_kernel<1>, dim3(nb), dim3(kBlock), 0, s, t, inv_scale,
                           cpos, grid, kanchor, dk, g, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Slabs of an aperture streamed through the device: acc = (first ? 0 : acc)
// plus the slab's partial images in split order, the order finalize_f64_kernel
// adds a resident pass's splits in; the last slab rounds the sum into img.
template <class PT>
__global__ __launch_bounds__(kBlock) void accumulate_f64_kernel(const PT *__restrict__ partial,
                                                                int nsplit, int npx, int first,
                                                                double2 *__restrict__ acc,
                                                                float2 *__restrict__ img)
{
    const int p = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    if (p >= npx) return;
    double sr = 0.0, si = 0.0;
    if (!first) {
        const double2 a = acc[p];
        sr = a.x;
        si = a.y;
    }
    add_splits(partial, nsplit, npx, p, sr, si);
    if (img) img[p] = make_float2(static_cast<float>(sr), static_cast<float>(si));
    else acc[p] = make_double2(sr, si);
}

hipError_t launch_accumulate_f64(const double2 *partial, int nsplit, int npx, int first,
                                 double2 *acc, float2 *img, hipStream_t s)
{
    if (npx <= 0) return hipSuccess;
    hipLaunchKernelGGL(accumulate_f64_kernel<double2>, dim3((npx + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, partial, nsplit, npx, first, acc, img);
    return hipGetLastError();
}

hipError_t launch_accumulate_p32(const float2 *partial, int nsplit, int npx, int first,
                                 double2 *acc, float2 *img, hipStream_t s)
{
    if (npx <= 0) return hipSuccess;
    hipLaunchKernelGGL(accumulate_f64_kernel<float2>, dim3((npx + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, partial, nsplit, npx, first, acc, img);
    return hipGetLastError();
}

hipError_t launch_finalize_f64(const double2 *partial, int nsplit, int npx, float2 *img,
                               hipStream_t s)
{
    hipLaunchKernelGGL(finalize_f64_kernel<double2>, dim3((npx + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, partial, nsplit, npx, img);
    return hipGetLastError();
}

hipError_t launch_finalize_p32(const float2 *partial, int nsplit, int npx, float2 *img,
                               hipStream_t s)
{
    hipLaunchKernelGGL(finalize_f64_kernel<float2>, dim3((npx + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, partial, nsplit, npx, img);
    return hipGetLastError();
}

} // namespace tdc
