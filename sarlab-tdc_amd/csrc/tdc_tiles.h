// tdc_tiles.h -- per-tile device bodies of the correlate-sum path, shared by
// the one-job kernels (tdc_kernels.hip) and the persistent serving kernel
// (tdc_serve.hip), so both run exactly the same arithmetic.
//
// A tile is up to kBlock * P consecutive pixels (lane l of the workgroup owns
// pixels pix0 + j * kBlock + l, j < P) times a chirp range [n0, n1). Every
// (chirp, sample) input is the same for all lanes of a wavefront: the plain
// form stages it in LDS and reads it into VGPRs, the vector precision form
// reads it with scalar loads, the matrix form with vector loads of
// pre-arranged MFMA operands. Pixel coordinates and accumulators live in
// VGPRs.
#pragma once

#include <hip/hip_runtime.h>

#include "tdc_internal.h"

namespace tdc {

// Read-only view in the constant address space: a wave-uniform load through
// it is always a scalar load, even in a kernel that also stores through other
// pointers the compiler cannot prove disjoint (the persistent serving kernel).
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using cptr = const __attribute__((address_space(4))) T *;
#else
template <class T> using cptr = const T *;
#endif
template <class T> __device__ __forceinline__ cptr<T> as_const(const T *p)
{
    return (cptr<T>) p;
}

// Word 3 of a CDNA buffer resource for untyped loads: data format 32 bits
// (a zero format would read as zero), no swizzle, no index stride
constexpr int kBufferRsrcWord3 = 0x00020000;

__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }
__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }

// |d| from |d|^2 in FP64: FP32 reciprocal-square-root seed and one FP64
// Newton step (relative error ~1e-13, i.e. ~1e-10 cycles of phase at metre
// ranges); zero stays zero.
__device__ __forceinline__ double range_from_sq(double r2)
{
    const double y0 = static_cast<double>(__builtin_amdgcn_rsqf(static_cast<float>(r2)));
    const double y1 = y0 * fma(-0.5 * r2 * y0, y0, 1.5);
    return r2 > 0.0 ? r2 * y1 : 0.0;
}

// |d| from |d|^2 in FP64 with three FP64 operations: r0 = r2 y0 from the FP32
// reciprocal-square-root seed y0, then one Newton step
// r0 + (r2 - r0^2) y0 / 2 (relative error ~1e-14). |d|^2 is first raised to
// at least 1e-30 m^2, so that zero (or a negative rounding of zero, at a
// pixel on the antenna's position) gives 1e-15 m, a phase of ~1e-12 cycles,
// with a finite seed.
__device__ __forceinline__ double range_newton(double r2)
{
    r2 = fmax(r2, 1e-30);
    const float y0f = __builtin_amdgcn_rsqf(static_cast<float>(r2));
    const double r0 = r2 * static_cast<double>(y0f);
    return fma(fma(-r0, r0, r2), static_cast<double>(0.5F * y0f), r0);
}

// x - floor(x) in FP64 (one instruction)
__device__ __forceinline__ double fract64(double x) { return __builtin_amdgcn_fract(x); }
// floor(x) as an int in one instruction (v_cvt_flr_i32_f32, which the compiler
// forms from a floor and a conversion only where it may assume no NaNs)
__device__ __forceinline__ int floor_to_int(float x)
{
    int i;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    return i;
}

// sin and cos of 2 pi x for x in [0, 1): reduction to the nearest quarter
// turn, |2 pi f| <= pi/4, then the Taylor series through a^9 (sin) and a^10
// (cos), whose truncation error there is below 2e-9.
__device__ __forceinline__ void sincos_rev(float x, float &s, float &c)
{
    const float q = __builtin_rintf(4.0F * x);
    const float f = fmaf(-0.25F, q, x);
    const float a = 6.28318530717958647692F * f;
    const float a2 = a * a;
    float ps = fmaf(a2, 2.75573192e-6F, -1.98412698e-4F);
    ps = fmaf(a2, ps, 8.33333333e-3F);
    ps = fmaf(a2, ps, -1.66666667e-1F);
    const float sa = fmaf(a * a2, ps, a);
    float pc = fmaf(a2, -2.75573192e-7F, 2.48015873e-5F);
    pc = fmaf(a2, pc, -1.38888889e-3F);
    pc = fmaf(a2, pc, 4.16666667e-2F);
    pc = fmaf(a2, pc, -0.5F);
    const float ca = fmaf(a2, pc, 1.0F);
    const int qi = static_cast<int>(q) & 3;
    // rotate (cos a, sin a) by qi quarter turns
    s = (qi & 1) ? ca : sa;
    c = (qi & 1) ? sa : ca;
    if (qi == 1 || qi == 2) c = -c;
    if (qi >= 2) s = -s;
}

// Error-free accumulation: acc + comp carries the running sum.
__device__ __forceinline__ void two_sum_acc(float &acc, float &comp, float x)
{
    const float s = acc + x;
    const float bp = s - acc;
    const float err = (acc - (s - bp)) + (x - bp);
    acc = s;
    comp += err;
}

// Raw sample loaders for the three storage formats; the address is uniform.
// round() gives an FP32 value the rounding of storage in the format, so that
// staging kernels can take raw samples straight from complex64 input.
struct RawC64 {
    using T = float2;
    static __device__ __forceinline__ void unpack(T v, float &re, float &im)
    {
        re = v.x;
        im = v.y;
    }
    static __device__ __forceinline__ void get(cptr<T> p, int i, float &re, float &im)
    {
        unpack(p[i], re, im);
    }
    static __device__ __forceinline__ void round(float &, float &) {}
};
struct RawF16 {
    using T = unsigned int;
    static __device__ __forceinline__ void unpack(T v, float &re, float &im)
    {
        re = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(v & 0xffffU)));
        im = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(v >> 16)));
    }
    static __device__ __forceinline__ void get(cptr<T> p, int i, float &re, float &im)
    {
        unpack(p[i], re, im);
        // convert once per sample and share the FP32 value across the lane's
        // pixels (left to itself the compiler folds the conversion into every
        // FMA as v_fma_mix_f32, which measured 40% slower)
        asm volatile("" : "+v"(re), "+v"(im));
    }
    static __device__ __forceinline__ void round(float &re, float &im)
    {
        re = static_cast<float>(static_cast<_Float16>(re));
        im = static_cast<float>(static_cast<_Float16>(im));
    }
};
struct RawBF16 {
    using T = unsigned int;
    static __device__ __forceinline__ void unpack(T v, float &re, float &im)
    {
        re = __uint_as_float(v << 16);
        im = __uint_as_float(v & 0xffff0000U);
    }
    static __device__ __forceinline__ void get(cptr<T> p, int i, float &re, float &im)
    {
        unpack(p[i], re, im);
    }
    // round to nearest even (finite inputs)
    static __device__ __forceinline__ void round(float &re, float &im)
    {
        unsigned int br = __float_as_uint(re), bi = __float_as_uint(im);
        br += 0x7fffU + ((br >> 16) & 1U);
        bi += 0x7fffU + ((bi >> 16) & 1U);
        re = __uint_as_float(br & 0xffff0000U);
        im = __uint_as_float(bi & 0xffff0000U);
    }
};

// ---------------------------------------------------------------------------
// Plain form (the drop-in): per term the direct reference phase
// theta = fract(k_m |pos - p|) in cycles, k_m = 2 (f0 + K t_m) / c, and a
// complex MAC of raw * exp(+j 2 pi theta) (reference tdckernels.cu:206-215,
// :270-272).
//
// Phase evaluation. k_m R reaches ~10^3 cycles, so forming it in FP32 leaves
// ~1e-4 cycles of rounding in every term. Instead the range R0 of a (pixel,
// chirp) pair is computed in FP64 (exact differences of the coordinates and
// an FP64 Newton square root), and every kAnchorBlock samples the phase of
// the block's anchor wavenumber k_b is reduced modulo one cycle in FP64,
// theta_b = fract(k_b R0). A term then adds dk_m R0 in FP32, dk_m = k_m - k_b
// tabulated per sample: theta_m = fract(fma(dk_m, R0, theta_b)), one FMA and
// a fract per term -- the operation count of fract(k_m R) -- with |dk_m R0|
// a few cycles, so ~1e-7 cycles of rounding.
//
// When the antenna moves within a chirp, the range of sample m is
// R0 + dR_m from the chirp's first position c: with delta = pos_m - c and
// d0 = p - c, |p - pos_m|^2 - R0^2 = delta . (delta - 2 d0) = num and
// dR = num / (|p - pos_m| + R0), formed from the small delta without
// cancellation; then theta_m = fract(theta_b + k_b dR + dk_m (R0 + dR)). For
// delta = 0 (stop-and-hop input) every extra term is an exact zero and the
// phase is bit-identical to the per-chirp form above. When the host has shown
// that the antenna moves little within a chirp against the shortest range
// (SERIES: k |delta|^3 / R^2 below 1e-7 cycles), dR comes from the first two
// terms of the expansion of the square root, dR = x - x^2 / (2 R0) with
// x = num / (2 R0), whose remainder x^3 / (2 R0^2) is below that bound: no
// square root and no reciprocal per term.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double range64(double px, double py, double pz, double cx, double cy,
                                          double cz)
{
    const double dx = px - cx, dy = py - cy, dz = pz - cz;
    return range_newton(fma(dz, dz, fma(dy, dy, dx * dx)));
}

// theta_b = fract(k_b R) in FP64, rounded to FP32
__device__ __forceinline__ float anchor_phase(double kb, double r)
{
    return static_cast<float>(fract64(kb * r));
}

// Per-term records of the plain form, motion within a chirp: the workgroup
// stages a chirp's records in LDS, kStageRecords at a time, and every lane
// reads each record (a broadcast ds_read): the records then reach the VALU as
// VGPR operands, which issue at twice the rate of SGPR operands on gfx950.
// Records: `rec` [npri][ns] PlainRec (raw * window, delta, dk_m), `cpos`
// [npri] float4 (x, y, z, 0) of each chirp's first sample, `kanchor`
// [ceil(ns / kAnchorBlock)] FP64 anchor wavenumbers.
constexpr int kStageRecords = 256;

template <int P, bool SERIES>
__device__ __forceinline__ void plain_tile(const PlainRec *__restrict__ rec,
                                           const float4 *__restrict__ cpos,
                                           const float *__restrict__ grid,
                                           const double *__restrict__ kanchor, int npx, int ns,
                                           int n0, int n1, int pix0, float2 *__restrict__ out)
{
    __shared__ float4 s_rec[2 * kStageRecords];
    double px[P], py[P], pz[P];
    float ar[P], ai[P];
    int pix[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        pix[j] = pix0 + j * kBlock + static_cast<int>(threadIdx.x);
        const size_t q = static_cast<size_t>(min(pix[j], npx - 1));
        px[j] = grid[3 * q];
        py[j] = grid[3 * q + 1];
        pz[j] = grid[3 * q + 2];
        ar[j] = 0.0F;
        ai[j] = 0.0F;
    }
    const cptr<double> ka = as_const(kanchor);
    for (int n = n0; n < n1; ++n) {
        const float4 c = as_const(cpos)[n];
        double rd[P];
        float r32[P], r0sq[P], hr[P], mx[P], my[P], mz[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            rd[j] = range64(px[j], py[j], pz[j], c.x, c.y, c.z);
            r32[j] = static_cast<float>(rd[j]);
            r0sq[j] = static_cast<float>(rd[j] * rd[j]);
            hr[j] = static_cast<float>(0.5 / rd[j]);
            // -2 d0, d0 = p - c
            mx[j] = static_cast<float>(-2.0 * (px[j] - c.x));
            my[j] = static_cast<float>(-2.0 * (py[j] - c.y));
            mz[j] = static_cast<float>(-2.0 * (pz[j] - c.z));
        }
        const float4 *src = reinterpret_cast<const float4 *>(rec + static_cast<size_t>(n) * ns);
        for (int m0 = 0; m0 < ns; m0 += kStageRecords) {
            const int cnt = min(kStageRecords, ns - m0);
            __syncthreads();
            for (int i = static_cast<int>(threadIdx.x); i < 2 * cnt; i += kBlock)
                s_rec[i] = src[2 * static_cast<size_t>(m0) + i];
            __syncthreads();
            for (int b0 = 0; b0 < cnt; b0 += kAnchorBlock) {
                const double kb = ka[(m0 + b0) / kAnchorBlock];
                const float kb32 = static_cast<float>(kb);
                float thb[P];
#pragma unroll
                for (int j = 0; j < P; ++j) thb[j] = anchor_phase(kb, rd[j]);
                const int b1 = min(cnt, b0 + kAnchorBlock);
#pragma unroll 4
                for (int e = b0; e < b1; ++e) {
                    // record fields: a = (rr, ri, delta x, delta y), b = (delta z, dk, pad, pad)
                    const float4 a = s_rec[2 * e], b = s_rec[2 * e + 1];
#pragma unroll
                    for (int j = 0; j < P; ++j) {
                        const float num =
                            fmaf(b.x, b.x + mz[j], fmaf(a.w, a.w + my[j], a.z * (a.z + mx[j])));
                        float dr;
                        if constexpr (SERIES) {
                            const float x = num * hr[j];
                            dr = fmaf(-(x * hr[j]), x, x);
                        } else {
                            const float rm = __builtin_amdgcn_sqrtf(r0sq[j] + num);
                            dr = num * __builtin_amdgcn_rcpf(rm + r32[j]);
                        }
                        const float th = __builtin_amdgcn_fractf(
                            fmaf(b.y, r32[j] + dr, fmaf(kb32, dr, thb[j])));
                        const float sn = sin_rev(th);
                        const float cs = cos_rev(th);
                        ar[j] = fmaf(a.x, cs, ar[j]);
                        ar[j] = fmaf(-a.y, sn, ar[j]);
                        ai[j] = fmaf(a.x, sn, ai[j]);
                        ai[j] = fmaf(a.y, cs, ai[j]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        if (pix[j] < npx) out[pix[j]] = make_float2(ar[j], ai[j]);
    }
}

// Stop-and-hop input: when the antenna position is the same for every sample
// of a chirp, the range is one value per (pixel, chirp) and a term's phase is
// fract(fma(dk_m, R0, theta_b)): the same operations plain_tile performs on
// such input, so the image is bit-identical to plain_tile's while the per-term
// range arithmetic leaves the loop. Records: `rec` [npri][ns] float4
// (rr, ri, dk_m, 0), `cpos` [npri] chirp positions (float4 (x, y, z, 0) or
// ChirpPos), `grid` FP32 or FP64 pixel coordinates.
constexpr int kTermStage = 512;

// Adds sum_m (rr + j ri) exp(+j 2 pi theta_m) over one chirp's records `row`
// to (sr, si), term by term in sample order, with the pair's FP64 range `rd`
// and its FP32 rounding `r32`. Every wave of the workgroup takes part in the
// staging; a wave with run == false (the same for all its lanes) only stages.
template <int P>
__device__ __forceinline__ void chirp_terms(const float4 *__restrict__ row, int ns,
                                            float4 *__restrict__ s_term,
                                            const double *__restrict__ kanchor,
                                            const double (&rd)[P], const float (&r32)[P],
                                            float (&sr)[P], float (&si)[P], bool run)
{
    const cptr<double> ka = as_const(kanchor);
    for (int m0 = 0; m0 < ns; m0 += kTermStage) {
        const int cnt = min(kTermStage, ns - m0);
        __syncthreads();
        for (int i = static_cast<int>(threadIdx.x); i < cnt; i += kBlock) s_term[i] = row[m0 + i];
        __syncthreads();
        if (!run) continue;
        for (int b0 = 0; b0 < cnt; b0 += kAnchorBlock) {
            const double kb = ka[(m0 + b0) / kAnchorBlock];
            float thb[P];
#pragma unroll
            for (int j = 0; j < P; ++j) thb[j] = anchor_phase(kb, rd[j]);
            const int b1 = min(cnt, b0 + kAnchorBlock);
#pragma unroll 4
            for (int e = b0; e < b1; ++e) {
                const float4 a = s_term[e];
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const float th = __builtin_amdgcn_fractf(fmaf(a.z, r32[j], thb[j]));
                    const float sn = sin_rev(th);
                    const float cs = cos_rev(th);
                    sr[j] = fmaf(a.x, cs, sr[j]);
                    sr[j] = fmaf(-a.y, sn, sr[j]);
                    si[j] = fmaf(a.x, sn, si[j]);
                    si[j] = fmaf(a.y, cs, si[j]);
                }
            }
        }
    }
}

template <int P, class GT, class CT>
__device__ __forceinline__ void range_tile(const float4 *__restrict__ rec,
                                           const CT *__restrict__ cpos,
                                           const GT *__restrict__ grid,
                                           const double *__restrict__ kanchor, int npx, int ns,
                                           int n0, int n1, int pix0, float2 *__restrict__ out)
{
    __shared__ float4 s_term[kTermStage];
    double px[P], py[P], pz[P];
    float ar[P], ai[P];
    int pix[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        pix[j] = pix0 + j * kBlock + static_cast<int>(threadIdx.x);
        const size_t q = static_cast<size_t>(min(pix[j], npx - 1));
        px[j] = grid[3 * q];
        py[j] = grid[3 * q + 1];
        pz[j] = grid[3 * q + 2];
        ar[j] = 0.0F;
        ai[j] = 0.0F;
    }
    for (int n = n0; n < n1; ++n) {
        const CT c = as_const(cpos)[n];
        double rd[P];
        float r32[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            rd[j] = range64(px[j], py[j], pz[j], c.x, c.y, c.z);
            r32[j] = static_cast<float>(rd[j]);
        }
        chirp_terms<P>(rec + static_cast<size_t>(n) * ns, ns, s_term, kanchor, rd, r32, ar, ai,
                       true);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        if (pix[j] < npx) out[pix[j]] = make_float2(ar[j], ai[j]);
    }
}

// Precision form over one tile: FP64 range per (pixel, chirp), FP64 anchor
// phase per 64-sample block, FP32 rotation recurrence in nested form inside
// the block, block sums accumulated per chirp, chirps accumulated with a
// two-sum compensation.
template <int P, class RAW>
__device__ __forceinline__ void precision_tile(
    const typename RAW::T *__restrict__ raw, const ChirpPos *__restrict__ cpos,
    const double *__restrict__ grid, const double *__restrict__ kanchor, double dk, int npx,
    int ns, int n0, int n1, int pix0, double2 *__restrict__ out)
{
    const int nblk = (ns + kAnchorBlock - 1) / kAnchorBlock;

    double px[P], py[P], pz[P];
    float accr[P], acci[P], cmpr[P], cmpi[P];
    int pix[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        pix[j] = pix0 + j * kBlock + static_cast<int>(threadIdx.x);
        const size_t q = static_cast<size_t>(min(pix[j], npx - 1));
        px[j] = grid[3 * q];
        py[j] = grid[3 * q + 1];
        pz[j] = grid[3 * q + 2];
        accr[j] = 0.0F;
        acci[j] = 0.0F;
        cmpr[j] = 0.0F;
        cmpi[j] = 0.0F;
    }

    for (int n = n0; n < n1; ++n) {
        const ChirpPos c = as_const(cpos)[n];
        double rng[P];
        float wr[P], wi[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const double dx = c.x - px[j], dy = c.y - py[j], dz = c.z - pz[j];
            rng[j] = range_from_sq(fma(dz, dz, fma(dy, dy, dx * dx)));
            double rot = dk * rng[j];
            rot -= floor(rot);
            sincos_rev(static_cast<float>(rot), wi[j], wr[j]);
        }
        const cptr<typename RAW::T> row = as_const(raw + static_cast<size_t>(n) * ns);
        float chr[P], chi[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            chr[j] = 0.0F;
            chi[j] = 0.0F;
        }
        for (int b = 0; b < nblk; ++b) {
            const int a = b * kAnchorBlock;
            const cptr<typename RAW::T> blk = row + a;
            float sr[P], si[P];
#pragma unroll
            for (int j = 0; j < P; ++j) {
                sr[j] = 0.0F;
                si[j] = 0.0F;
            }
            const int len = min(kAnchorBlock, ns - a);
            if (len == kAnchorBlock) {
#pragma unroll
                for (int i = kAnchorBlock - 1; i >= 0; --i) {
                    float vr, vi;
                    RAW::get(blk, i, vr, vi);
#pragma unroll
                    for (int j = 0; j < P; ++j) {
                        const float nr = fmaf(sr[j], wr[j], fmaf(-si[j], wi[j], vr));
                        const float ni = fmaf(sr[j], wi[j], fmaf(si[j], wr[j], vi));
                        sr[j] = nr;
                        si[j] = ni;
                    }
                }
            } else {
                for (int i = len - 1; i >= 0; --i) {
                    float vr, vi;
                    RAW::get(blk, i, vr, vi);
#pragma unroll
                    for (int j = 0; j < P; ++j) {
                        const float nr = fmaf(sr[j], wr[j], fmaf(-si[j], wi[j], vr));
                        const float ni = fmaf(sr[j], wi[j], fmaf(si[j], wr[j], vi));
                        sr[j] = nr;
                        si[j] = ni;
                    }
                }
            }
            const double ka = as_const(kanchor)[b];
#pragma unroll
            for (int j = 0; j < P; ++j) {
                double th = ka * rng[j];
                th -= floor(th);
                const float tf = static_cast<float>(th);
                const float zr = cos_rev(tf), zi = sin_rev(tf);
                chr[j] += fmaf(zr, sr[j], -(zi * si[j]));
                chi[j] += fmaf(zr, si[j], zi * sr[j]);
            }
        }
        // compensated accumulation over chirps
#pragma unroll
        for (int j = 0; j < P; ++j) {
            two_sum_acc(accr[j], cmpr[j], chr[j]);
            two_sum_acc(acci[j], cmpi[j], chi[j]);
        }
    }

#pragma unroll
    for (int j = 0; j < P; ++j) {
        if (pix[j] < npx) {
            out[pix[j]] = make_double2(static_cast<double>(accr[j]) + static_cast<double>(cmpr[j]),
                                     static_cast<double>(acci[j]) + static_cast<double>(cmpi[j]));
        }
    }
}

// ---------------------------------------------------------------------------
// Matrix engine of the precision form.
//
// With i = 16a + c inside a 128-sample block b, w = exp(j 2 pi dk R) and
// u = w^16,
//   S_b = sum_i raw[128b + i] w^i = sum_a u^a T_{b,a},
//   T_{b,a} = sum_{c<16} raw[128b + 16a + c] w^c,
// and T is a matrix product: rows (block, a, re/im) from the chirp's raw
// samples (the A operand, the same for every pixel), columns pixels, and the
// contraction over (c, re/im of w^c) (the B operand, per pixel). Two chained
// v_mfma_f32_32x32x16_f16 (k-slices c < 8 and c >= 8) compute T for two
// blocks (256 samples) and 32 pixels. FP16 operands are split hi + lo (each
// value = hi + lo to ~2^-22), so the product keeps FP32-level accuracy:
// hi*hi + hi*lo (+ lo*hi for c64 raw). The VALU then runs the 8-step nested
// form over a with u, and rotates each block sum by its FP64 anchor phase, as
// the one-sample form does every 64 samples.
//
// Lane layout (wave64; l = lane, g = l / 32): lane l owns pixel wave_base + l
// and computes its w^0..w^15; the B operand of pixel half h (pixels
// wave_base + 32h + (l % 32)) and k-slice s holds Re w^(8s+j) in lanes g = 0
// and Im w^(8s+j) in lanes g = 1. Accumulator register r of lane l holds
// T_{2q+g, r/2} of pixel half h, real part for even r, q the tile set. So lane
// group g sums the even (g = 0) or odd (g = 1) blocks of both pixel halves;
// the two groups are added at the end.
// ---------------------------------------------------------------------------

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef unsigned int uintx4_t __attribute__((ext_vector_type(4)));

// Lanes 0..31 of the first result hold x of lanes 0..31, lanes 32..63 hold y
// of lanes 0..31; the second result holds x then y of lanes 32..63.
__device__ __forceinline__ void swap_halves(unsigned int x, unsigned int y, unsigned int &lo,
                                            unsigned int &hi)
{
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    lo = r[0];
    hi = r[1];
}
__device__ __forceinline__ void swap_halves(float x, float y, float &lo, float &hi)
{
    unsigned int a, b;
    swap_halves(__float_as_uint(x), __float_as_uint(y), a, b);
    lo = __uint_as_float(a);
    hi = __uint_as_float(b);
}
__device__ __forceinline__ void swap_halves(double x, double y, double &lo, double &hi)
{
    const uint2 xa = __builtin_bit_cast(uint2, x), ya = __builtin_bit_cast(uint2, y);
    uint2 l, h;
    swap_halves(xa.x, ya.x, l.x, h.x);
    swap_halves(xa.y, ya.y, l.y, h.y);
    lo = __builtin_bit_cast(double, l);
    hi = __builtin_bit_cast(double, h);
}
__device__ __forceinline__ void swap_halves(half8_t x, half8_t y, half8_t &lo, half8_t &hi)
{
    const uintx4_t xa = __builtin_bit_cast(uintx4_t, x), ya = __builtin_bit_cast(uintx4_t, y);
    uintx4_t l, h;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        unsigned int a, b;
        swap_halves(xa[k], ya[k], a, b);
        l[k] = a;
        h[k] = b;
    }
    lo = __builtin_bit_cast(half8_t, l);
    hi = __builtin_bit_cast(half8_t, h);
}

// FP16 hi and lo parts of two FP32 values, each packed as (first, second):
// hi = f16(v), lo = f16(v - hi), both rounded to nearest even; the
// subtraction is exact in FP32 and v_fma_mix rounds it to FP16 once.
__device__ __forceinline__ void split_f16x2(float v0, float v1, unsigned int &hi, unsigned int &lo)
{
    asm("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(hi) : "v"(v0), "v"(v1));
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=&v"(lo) : "v"(hi), "v"(v0));
    asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "+v"(lo)
        : "v"(hi), "v"(v1));
}

__device__ __forceinline__ void cmul(float ar, float ai, float br, float bi, float &cr, float &ci)
{
    cr = fmaf(ar, br, -(ai * bi));
    ci = fmaf(ar, bi, ai * br);
}

// Accumulate the 8 group sums of one block with the nested form in u and
// rotate the block sum by the anchor exp(j 2 pi theta).
__device__ __forceinline__ void matrix_block(const floatx16_t &d, float ur, float ui, float theta,
                                             float &chr, float &chi)
{
    float sr = d[14], si = d[15];
#pragma unroll
    for (int a = 6; a >= 0; --a) {
        const float nr = fmaf(sr, ur, fmaf(-si, ui, d[2 * a]));
        const float ni = fmaf(sr, ui, fmaf(si, ur, d[2 * a + 1]));
        sr = nr;
        si = ni;
    }
    const float zr = cos_rev(theta), zi = sin_rev(theta);
    chr = fmaf(zr, sr, fmaf(-zi, si, chr));
    chi = fmaf(zr, si, fmaf(zi, sr, chi));
}

__device__ __forceinline__ floatx16_t mfma16(const uintx4_t &a, const half8_t &b,
                                             const floatx16_t &c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8_t, a), b, c, 0, 0, 0);
}

// One tile: kBlock consecutive pixels from pix0 (one per thread) times the
// chirp range [n0, n1). atab holds, per (chirp, tile set), the A tiles of
// k-slices 0 and 1, each PARTS parts of 64 lanes x 16 bytes; inv_scale undoes
// the per-chirp power-of-two scale of the A tiles; kanchor holds one
// wavenumber per 64 samples.
template <int PARTS>
__device__ __forceinline__ void matrix_tile(const uintx4_t *__restrict__ atab,
                                            const float *__restrict__ inv_scale,
                                            const ChirpPos *__restrict__ cpos,
                                            const double *__restrict__ grid,
                                            const double *__restrict__ kanchor, double dk, int npx,
                                            int nq, int n0, int n1, int pix0,
                                            double2 *__restrict__ out)
{
    constexpr int kTiles = 2 * PARTS; // A tiles per tile set
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const bool upper = lane >= 32;
    const int pix = pix0 + static_cast<int>(threadIdx.x);
    const size_t q = static_cast<size_t>(min(pix, npx - 1));
    const double px = grid[3 * q], py = grid[3 * q + 1], pz = grid[3 * q + 2];
    const double dk16 = 16.0 * dk;
    double accr[2] = {0.0, 0.0}, acci[2] = {0.0, 0.0};
    const cptr<double> ka = as_const(kanchor);
    const uintx4_t *tp = atab + static_cast<size_t>(max(n0, 0)) * nq * kTiles * 64 + lane;
    uintx4_t nxt[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; ++t) nxt[t] = uintx4_t{0U, 0U, 0U, 0U};
    if (n0 < n1) {
#pragma unroll
        for (int t = 0; t < kTiles; ++t) nxt[t] = tp[64 * t];
    }

    for (int n = n0; n < n1; ++n) {
        const ChirpPos c = as_const(cpos)[n];
        const double dx = c.x - px, dy = c.y - py, dz = c.z - pz;
        const double rng = range_from_sq(fma(dz, dz, fma(dy, dy, dx * dx)));
        // w and u = w^16 straight from their FP64 fractions, with the hardware
        // sin/cos (within 1.2e-7 rad and 1.2e-7 of unit magnitude over a
        // turn, the size of the FP32 rounding of the fraction itself)
        const float xw = static_cast<float>(fract64(dk * rng));
        const float xu = static_cast<float>(fract64(dk16 * rng));
        const float wr = cos_rev(xw), wi = sin_rev(xw);
        const float ur0 = cos_rev(xu), ui0 = sin_rev(xu);

        // w^c, c = 0..15, split into FP16 hi + lo
        float pr[kGroup], pi[kGroup];
        pr[0] = 1.0F;
        pi[0] = 0.0F;
        pr[1] = wr;
        pi[1] = wi;
#pragma unroll
        for (int k = 2; k < kGroup; ++k) {
            // w^k from two lower powers, each of at most k/2 + 1 factors
            const int a = k / 2, b = k - k / 2;
            cmul(pr[a], pi[a], pr[b], pi[b], pr[k], pi[k]);
        }
        half8_t bh[2][2], bl[2][2]; // [k-slice][pixel half]
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            uintx4_t rh, ih, rl, il;
#pragma unroll
            for (int k = 0; k < kSlice / 2; ++k) {
                const int e = kSlice * s + 2 * k;
                unsigned int h, l;
                split_f16x2(pr[e], pr[e + 1], h, l);
                rh[k] = h;
                rl[k] = l;
                split_f16x2(pi[e], pi[e + 1], h, l);
                ih[k] = h;
                il[k] = l;
            }
            swap_halves(__builtin_bit_cast(half8_t, rh), __builtin_bit_cast(half8_t, ih),
                        bh[s][0], bh[s][1]);
            swap_halves(__builtin_bit_cast(half8_t, rl), __builtin_bit_cast(half8_t, il),
                        bl[s][0], bl[s][1]);
        }
        float ur[2], ui[2];
        double rh2[2];
        swap_halves(ur0, ur0, ur[0], ur[1]);
        swap_halves(ui0, ui0, ui[0], ui[1]);
        swap_halves(rng, rng, rh2[0], rh2[1]);

        float chr[2] = {0.0F, 0.0F}, chi[2] = {0.0F, 0.0F};
        for (int p = 0; p < nq; ++p) {
            uintx4_t cur[kTiles];
#pragma unroll
            for (int t = 0; t < kTiles; ++t) cur[t] = nxt[t];
            // prefetch the next (chirp, tile set) of the tile; the tables are
            // contiguous over chirps, the last one reloads itself
            if (n + 1 < n1 || p + 1 < nq) tp += kTiles * 64;
#pragma unroll
            for (int t = 0; t < kTiles; ++t) nxt[t] = tp[64 * t];
            floatx16_t d0 = {}, d1 = {};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const uintx4_t &ah = cur[s * PARTS];
                d0 = mfma16(ah, bh[s][0], d0);
                d1 = mfma16(ah, bh[s][1], d1);
                d0 = mfma16(ah, bl[s][0], d0);
                d1 = mfma16(ah, bl[s][1], d1);
                if (PARTS == 2) {
                    const uintx4_t &al = cur[s * PARTS + 1];
                    d0 = mfma16(al, bh[s][0], d0);
                    d1 = mfma16(al, bh[s][1], d1);
                }
            }
            // anchor wavenumber of this lane group's block (128 samples = two
            // entries of the 64-sample table)
            const double ke = ka[4 * p], ko = ka[4 * p + 2];
            const double kb = upper ? ko : ke;
            matrix_block(d0, ur[0], ui[0], static_cast<float>(fract64(kb * rh2[0])), chr[0], chi[0]);
            matrix_block(d1, ur[1], ui[1], static_cast<float>(fract64(kb * rh2[1])), chr[1], chi[1]);
        }
        const float is = as_const(inv_scale)[n];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            accr[h] += static_cast<double>(chr[h] * is);
            acci[h] += static_cast<double>(chi[h] * is);
        }
    }

    // even-block and odd-block sums of the lane's own pixel
    double ar0, ar1, ai0, ai1;
    swap_halves(accr[0], accr[1], ar0, ar1);
    swap_halves(acci[0], acci[1], ai0, ai1);
    if (pix < npx) out[pix] = make_double2(ar0 + ar1, ai0 + ai1);
}

// Range-profile engine over one tile (one pixel per thread, chirps n0..n1):
// per (pixel, chirp) the range R, the bin position u = frac(dk R) M split
// into an integer bin f and a fraction g, the TAPS table entries around u
// weighted by the Gaussian kernel (two exp2 per pair, two multiplies per
// tap), and the anchor exp(j 2 pi k_h R); chirps accumulate in FP32 blocks
// of kProfileBlock chirps and the blocks in FP64, and the tile's partial
// image is stored rounded to FP32 (its splits are added in FP64). R, u and the anchor phase
// are evaluated in FP64 for the first chirp of each block and, with pp.incr,
// advanced by FP32 increments for the block's other chirps (else in FP64 for
// every chirp).
namespace profile {
// A pixel's bin position u = frac(dk R) M has integer part f and fraction g;
// tap t (0 .. TAPS - 1) is bin f - lead + t, lead = TAPS / 2 - 1, at distance
// g + lead - t from u, and weighs exp(-b (g + lead - t)^2) = up0 e1^t c[t] with
// up0 = exp(-b g^2 - 2 lead b g), e1 = exp(2 b g), c[t] = exp(-b (lead - t)^2)
// (fast Gaussian gridding: two exponentials per pixel and chirp, then one
// product per tap); b = k pi / (TAPS / 2).
constexpr int kProfileBlock = 8;
// Constants of the built kernels (taps, oversampling), generated from
// b = k pi / (taps / 2) with the k of profile_bfactor: kE0 = -b log2(e),
// kE1 = 2 b log2(e), kC[t] = exp(-b (taps / 2 - 1 - t)^2).
template <int TAPS, int OVER> struct Kernel;
template <> struct Kernel<8, 2> { // k = 0.80
    static constexpr float kE0 = -0.9064720283654387F;
    static constexpr float kE1 = 1.8129440567308774F;
    static constexpr float kC[8] = {0.003500439396667034F, 0.08100259215794314F,
                                    0.5334880910911033F,   1.0F,
                                    0.5334880910911033F,   0.08100259215794314F,
                                    0.003500439396667034F, 4.305223158055476e-05F};
};
template <> struct Kernel<10, 2> { // k = 0.80
    static constexpr float kE0 = -0.7251776226923511F;
    static constexpr float kE1 = 1.4503552453847022F;
    static constexpr float kC[10] = {0.00032151151668886733F, 0.010846710538160161F,
                                     0.133905721399763F,      0.6049225627642709F,
                                     1.0F,                    0.6049225627642709F,
                                     0.133905721399763F,      0.010846710538160161F,
                                     0.00032151151668886733F, 3.487342356208991e-06F};
};
template <> struct Kernel<12, 2> { // k = 0.79
    static constexpr float kE0 = -0.5967607520072472F;
    static constexpr float kE1 = 1.1935215040144944F;
    static constexpr float kC[12] = {3.227957702166608e-05F, 0.0013357150352831012F,
                                     0.02416651468988259F,   0.19117383558744308F,
                                     0.6612369490950528F,    1.0F,
                                     0.6612369490950528F,    0.19117383558744308F,
                                     0.02416651468988259F,   0.0013357150352831012F,
                                     3.227957702166608e-05F, 3.410798652928483e-07F};
};
template <> struct Kernel<6, 4> { // k = 0.87
    static constexpr float kE0 = -1.3143844411298862F;
    static constexpr float kE1 = 2.6287688822597723F;
    static constexpr float kC[6] = {0.02614107418320684F, 0.40209702267679975F,
                                    1.0F,                 0.40209702267679975F,
                                    0.02614107418320684F, 0.0002747753163046619F};
};
template <> struct Kernel<8, 4> { // k = 0.88
    static constexpr float kE0 = -0.9971192312019825F;
    static constexpr float kE1 = 1.994238462403965F;
    static constexpr float kC[8] = {0.0019885422346150906F, 0.0630011981095824F,
                                    0.5009993958482307F,    1.0F,
                                    0.5009993958482307F,    0.0630011981095824F,
                                    0.0019885422346150906F, 1.5754159369011635e-05F};
};
} // namespace profile

// Taps from a window of the table row staged by the wave itself: a wave's
// 64 pixels are close together, so their first taps usually lie within
// (64 - TAPS) / 2 bins of lane 0's. Each lane then loads one entry of the
// 64-entry window starting that far below lane 0's first tap (one coalesced
// load) into the wave's own LDS slot, and reads its taps from there (LDS
// delivers twice the bytes per clock of the vector L1, which the gathered
// taps would otherwise saturate). A wave with a lane outside the window
// gathers from the table directly.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int TAPS, int OVER>
__device__ __forceinline__ void profile_tile(const float2 *__restrict__ table,
                                             const ChirpPos *__restrict__ cpos,
                                             const double *__restrict__ grid,
                                             const ProfileParams &pp, int npx, int n0, int n1,
                                             int pix0, float2 *__restrict__ out)
{
    __shared__ float2 s_win[kBlock / 64][64];
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const unsigned lane_bytes = static_cast<unsigned>(lane) * static_cast<unsigned>(sizeof(float2));
    float2 *win = s_win[threadIdx.x >> 6];
    const int pix = pix0 + static_cast<int>(threadIdx.x);
    const size_t q = static_cast<size_t>(min(pix, npx - 1));
    // |pos - p|^2 = |pos'|^2 - 2 pos'.p' + |p'|^2 with both taken from an
    // origin inside the scene (no cancellation beyond the scene's own size):
    // four FP64 operations per chirp instead of six
    const double px = grid[3 * q] - pp.ox, py = grid[3 * q + 1] - pp.oy, pz = grid[3 * q + 2] - pp.oz;
    const double pp2 = fma(pz, pz, fma(py, py, px * px));
    const double qx = -2.0 * px, qy = -2.0 * py, qz = -2.0 * pz;
    const int mask = pp.M - 1;
    // FP32 operands of the incremental range (pp.incr): -p', |dk| M and k_h
    // (their rounding changes a step's bin and phase by a relative 2^-24 of
    // |dk| M dR and k_h dR, inside profile_incremental_ok's bound)
    const float qx32 = static_cast<float>(0.5 * qx), qy32 = static_cast<float>(0.5 * qy),
                qz32 = static_cast<float>(0.5 * qz);
    const float dkm32 = static_cast<float>(pp.dkm);
    const float kh32 = static_cast<float>(pp.kh);
    // Within a block the rows and chirp records are addressed by 32-bit byte
    // offsets from the block's first (a row is at most (kProfileMaxM +
    // kProfilePad) 8 bytes), the rows through one buffer resource per block:
    // row n starts (n - nb) Mp 8 bytes past row nb, its position increment
    // kProfileDeltaEntry 8 bytes past its M bins
    const unsigned row_bytes = static_cast<unsigned>(pp.Mp) * static_cast<unsigned>(sizeof(float2));
    const unsigned delta_bytes =
        static_cast<unsigned>(pp.M + kProfileDeltaEntry) * static_cast<unsigned>(sizeof(float2));
    // all ones when the block's other chirps advance by FP32 increments, else
    // zero: an integer kept opaque, so that the per-chirp choice stays one
    // scalar AND and compare instead of being folded into lane-mask logic
    unsigned incr_mask = __builtin_amdgcn_readfirstlane(pp.incr ? ~0U : 0U);
    asm volatile("" : "+s"(incr_mask));
    // The chirps in order, each block of kProfileBlock from n0 with its first
    // chirp exact; chirp n + 1's state and the load of its window are issued
    // while chirp n's taps are summed. The block's rows and chirp records
    // are addressed from its first: brow, its buffer resource, cbytes.
    double accr = 0.0, acci = 0.0;
    if (n0 < n1) {
        float s = 0.0F, is = 0.0F, g = 0.0F, a = 0.0F, br = 0.0F, bi = 0.0F;
        unsigned f = 0;
        const float2 *brow = table + static_cast<size_t>(n0) * pp.Mp;
        auto tab = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(brow),
                                                     static_cast<short>(0), 0x7fffffff,
                                                     kBufferRsrcWord3);
        const char *cbytes = reinterpret_cast<const char *>(cpos + n0);
        // the pixel's state at the chirp: FP32 range s and 1 / s, bin f and
        // fraction g of its position, anchor phase a = frac(k_h R); within a
        // block f is kept modulo 2^32 (a multiple of M) and masked to the row
        // where a row is indexed with it
        unsigned row0 = 0, k = 0; // the chirp's row offset and place in its block
        // the state in FP64: a block's first chirp (k = 0), or every chirp
        // without pp.incr
        auto exact = [&]() {
            const ChirpPos c = as_const(reinterpret_cast<const ChirpPos *>(
                cbytes + (k & ~incr_mask) * static_cast<unsigned>(sizeof(ChirpPos))))[0];
            const double rng = range_newton(fma(c.z, qz, fma(c.y, qy, fma(c.x, qx, c.pad + pp2))));
            // bin position u = x mod M, x = |dk| M R >= 0: integer part f (the
            // truncation of x, exact while x < 2^31, i.e. R B < ~8e16 m Hz at
            // fourfold oversampling) and fraction g; down-chirps read reversed
            // rows, so x is never negative
            const double x = pp.dkm * rng;
            g = static_cast<float>(fract64(x));
            // indices in unsigned arithmetic: f < M, so row[f .. f + TAPS - 1]
            // lies in the row; window indices are masked to the row
            f = static_cast<unsigned>(static_cast<int>(x)) & static_cast<unsigned>(mask);
            a = static_cast<float>(fract64(pp.kh * rng));
            // (read by the next chirp's increment only when pp.incr)
            s = static_cast<float>(rng);
            is = __builtin_amdgcn_rcpf(s);
        };
        // the state advanced by one chirp within a block: delta / 2 = (R_n^2
        // - R_{n-1}^2) / 2 from the chirp's position increment (the row's last
        // entries), dR = e0 (1 - u / 2 + u^2 / 2), e0 = delta / (2 R), u = e0 /
        // R (the series of sqrt(R^2 + delta) - R), then s, the bin position
        // and the anchor phase advanced by dR (bounds: profile_incremental_ok)
        auto advance = [&]() {
            const float4 d = as_const(reinterpret_cast<const float4 *>(
                reinterpret_cast<const char *>(brow) + (row0 + delta_bytes)))[0];
            const float hdelta = fmaf(qz32, d.z, fmaf(qy32, d.y, fmaf(qx32, d.x, d.w)));
            const float e0 = hdelta * is;
            const float u = e0 * is;
            const float dr = e0 * fmaf(u, fmaf(0.5F, u, -0.5F), 1.0F);
            s += dr;
            is = __builtin_amdgcn_rcpf(s);
            const float gx = fmaf(dkm32, dr, g);
            f += static_cast<unsigned>(floor_to_int(gx));
            g = __builtin_amdgcn_fractf(gx);
            a = __builtin_amdgcn_fractf(fmaf(kh32, dr, a));
        };
        // the wave's window of the chirp's row: entries lo .. lo + 63 modulo
        // M, lo (64 - TAPS) / 2 bins below lane 0's first tap (from bin lo mod
        // M on, the row's wrap-around entries cover the window's end), by a
        // buffer load: the block's rows as a scalar resource, the window's
        // start as its scalar offset, the lane's byte offset in a
        // loop-invariant register
        constexpr unsigned kReach = (64 - TAPS) / 2;
        auto window = [&](unsigned &lo) {
            lo = __builtin_amdgcn_readfirstlane(f) - kReach;
            // (lo kept opaque, so that the window test's off also addresses
            // the window instead of being formed again from lane 0's bin)
            asm volatile("" : "+s"(lo));
            return __builtin_bit_cast(
                float2, __builtin_amdgcn_raw_buffer_load_b64(
                            tab, lane_bytes,
                            static_cast<int>(row0 + (lo & static_cast<unsigned>(mask)) *
                                                        static_cast<unsigned>(sizeof(float2))),
                            0));
        };
        exact();
        unsigned lo;
        float2 e = window(lo);
        for (int n = n0; n < n1; ++n) {
            // chirp n's taps: from its window (loaded during chirp n - 1)
            // through the wave's LDS slot, or, with a lane outside the
            // window, gathered from the row
            const unsigned off = f - lo;
            float2 v[TAPS];
            if (__all(off <= 64U - TAPS)) {
                wave_lds_sync(); // the previous chirp's reads of the slot are done
                win[lane] = e;
                wave_lds_sync();
#pragma unroll
                for (int t = 0; t < TAPS; ++t) v[t] = win[off + t];
            } else {
                const float2 *r = reinterpret_cast<const float2 *>(
                                      reinterpret_cast<const char *>(brow) + row0) +
                                  (f & static_cast<unsigned>(mask));
#pragma unroll
                for (int t = 0; t < TAPS; ++t) v[t] = r[t];
                // (waited for here, so that the sums below wait for no
                // vector-memory load and the next window's stays in flight)
#pragma unroll
                for (int t = 0; t < TAPS; ++t) asm volatile("" ::"v"(v[t].x), "v"(v[t].y));
            }
            // chirp n's weights and anchor rotation, formed before its state
            // moves on: up0 = exp2(g (kE0 g - lead kE1)), e1 = exp2(kE1 g)
            // (for g in [0, 1) and every built kernel up0 lies in (0.0079, 1]
            // and up0 e1^(TAPS - 1) below 126, far inside FP32's range)
            using K = profile::Kernel<TAPS, OVER>;
            constexpr int kLead = TAPS / 2 - 1;
            float up = __builtin_amdgcn_exp2f(g * fmaf(K::kE0, g, -kLead * K::kE1));
            const float e1 = __builtin_amdgcn_exp2f(K::kE1 * g);
            const float zr = cos_rev(a), zi = sin_rev(a);
            asm volatile("" ::"v"(up), "v"(e1), "v"(zr), "v"(zi));
            // chirp n + 1: its state and its window's load, in flight while
            // chirp n's taps are summed
            k = k + 1 == static_cast<unsigned>(profile::kProfileBlock) ? 0U : k + 1;
            const bool last = n + 1 == n1;
            if (!last) {
                if (k == 0) {
                    brow += static_cast<size_t>(profile::kProfileBlock) * pp.Mp;
                    tab = __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(brow),
                                                            static_cast<short>(0), 0x7fffffff,
                                                            kBufferRsrcWord3);
                    cbytes += profile::kProfileBlock * sizeof(ChirpPos);
                    row0 = 0;
                } else {
                    row0 += row_bytes;
                }
                if ((k & incr_mask) == 0U)
                    exact();
                else
                    advance();
                e = window(lo);
            }
            float ar = 0.0F, ai = 0.0F;
#pragma unroll
            for (int t = 0; t < TAPS; ++t) {
                const float w = up * K::kC[t];
                ar = fmaf(w, v[t].x, ar);
                ai = fmaf(w, v[t].y, ai);
                up *= e1;
            }
            br = fmaf(zr, ar, fmaf(-zi, ai, br));
            bi = fmaf(zr, ai, fmaf(zi, ar, bi));
            // chirp contributions are summed in FP32 over a block and the
            // blocks in FP64
            if (k == 0 || last) {
                accr += static_cast<double>(br);
                acci += static_cast<double>(bi);
                br = 0.0F;
                bi = 0.0F;
            }
        }
    }
    if (pix < npx) out[pix] = make_float2(static_cast<float>(accr), static_cast<float>(acci));
}

} // namespace tdc
