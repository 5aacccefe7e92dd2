// tdc_tiles.h -- per-tile device bodies of the correlate-sum path, shared by
// the one-job kernels (tdc_kernels.hip) and the persistent serving kernel
// (tdc_serve.hip), so both run exactly the same arithmetic.
//
// A tile is up to kBlock * P consecutive pixels (lane l of the workgroup owns
// pixels pix0 + j * kBlock + l, j < P) times a chirp range [n0, n1). Every
// (chirp, sample) input is the same for all lanes of a wavefront and is read
// with scalar loads; pixel coordinates and accumulators live in VGPRs.
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
struct RawC64 {
    using T = float2;
    static __device__ __forceinline__ void get(cptr<T> p, int i, float &re, float &im)
    {
        const float2 v = p[i];
        re = v.x;
        im = v.y;
    }
};
struct RawF16 {
    using T = unsigned int;
    static __device__ __forceinline__ void get(cptr<T> p, int i, float &re, float &im)
    {
        const unsigned int v = p[i];
        re = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(v & 0xffffU)));
        im = static_cast<float>(__builtin_bit_cast(_Float16, static_cast<unsigned short>(v >> 16)));
        // convert once per sample and share the FP32 value across the lane's
        // pixels (left to itself the compiler folds the conversion into every
        // FMA as v_fma_mix_f32, which measured 40% slower)
        asm volatile("" : "+v"(re), "+v"(im));
    }
};
struct RawBF16 {
    using T = unsigned int;
    static __device__ __forceinline__ void get(cptr<T> p, int i, float &re, float &im)
    {
        const unsigned int v = p[i];
        re = __uint_as_float(v << 16);
        im = __uint_as_float(v & 0xffff0000U);
    }
};

// Plain form over one tile, Doppler window identically one: per term the
// direct phase theta = fract(k_m |pos - p|) in cycles and a complex MAC of
// raw * exp(+j 2 pi theta) (reference tdckernels.cu:206-215, :270-272).
template <int P>
__device__ __forceinline__ void plain_tile(const PlainRec *__restrict__ rec,
                                           const float *__restrict__ grid, int npx, int ns,
                                           int n0, int n1, int pix0, float2 *__restrict__ out)
{

    float px[P], py[P], pz[P], ar[P], ai[P];
    int pix[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        pix[j] = pix0 + j * kBlock + static_cast<int>(threadIdx.x);
        const int q = min(pix[j], npx - 1);
        px[j] = grid[3 * q];
        py[j] = grid[3 * q + 1];
        pz[j] = grid[3 * q + 2];
        ar[j] = 0.0F;
        ai[j] = 0.0F;
    }

    const cptr<PlainRec> r = as_const(rec + static_cast<size_t>(max(n0, 0)) * ns);
    const int nterms = (n1 > n0) ? (n1 - n0) * ns : 0;
#pragma unroll 4
    for (int e = 0; e < nterms; ++e) {
        const PlainRec t = r[e];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const float dx = px[j] - t.x;
            const float dy = py[j] - t.y;
            const float dz = pz[j] - t.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const float th = __builtin_amdgcn_fractf(__builtin_amdgcn_sqrtf(r2) * t.k);
            const float sn = sin_rev(th);
            const float cs = cos_rev(th);
            ar[j] = fmaf(t.rr, cs, ar[j]);
            ar[j] = fmaf(-t.ri, sn, ar[j]);
            ai[j] = fmaf(t.rr, sn, ai[j]);
            ai[j] = fmaf(t.ri, cs, ai[j]);
        }
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
        const int q = min(pix[j], npx - 1);
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

} // namespace tdc
