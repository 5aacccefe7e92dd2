// scene.cpp -- FMCW beat-signal simulator for point-scatterer scenes (host only).
//
// Produces the raw data a linear-FMCW radar would record from a set of point
// scatterers, in the sign convention the TDC correlator focuses:
//   raw[n, m] = sum_k a_k * exp(-j 2 pi * 2 (f0 + K t_m) |track[n] - s_k| / c)
// (the correlator multiplies by exp(+j 2 pi * 2 f R / c); reference
// tdc-lib/src/tdckernels.cu:206-215 and :270-272).
//
// Evaluated in FP64. For uniformly spaced sample times the per-sample phasor
// advances by a constant FP64 rotation and is re-anchored with a direct
// sincos every 32 samples, which keeps the error near 1e-14; otherwise every
// sample is evaluated directly.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr int kAnchor = 32;

bool uniform_times(const double *t, int ns, double *dt)
{
    if (ns < 2) {
        *dt = 0.0;
        return true;
    }
    double step = (t[ns - 1] - t[0]) / (ns - 1);
    double tol = std::fabs(step) * 1e-9 + 1e-300;
    for (int m = 0; m < ns; ++m) {
        if (std::fabs(t[m] - (t[0] + m * step)) > tol) return false;
    }
    *dt = step;
    return true;
}

void sim_rows(const double *track, int n0, int n1, int ns, const double *t, double f0, double k_rate,
              double c, const double *scat, const double *amp, int nk, double *out, bool uniform,
              double dt)
{
    for (int n = n0; n < n1; ++n) {
        double *row = out + static_cast<size_t>(n) * ns * 2;
        std::memset(row, 0, sizeof(double) * 2 * ns);
        const double *tp = track + static_cast<size_t>(n) * 3;
        for (int k = 0; k < nk; ++k) {
            const double *s = scat + static_cast<size_t>(k) * 3;
            double dx = tp[0] - s[0], dy = tp[1] - s[1], dz = tp[2] - s[2];
            double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            double ar = amp[2 * k], ai = amp[2 * k + 1];
            double two_r_over_c = 2.0 * r / c;
            if (uniform) {
                // per-sample rotation exp(-j 2 pi * 2 K dt R / c)
                double dcyc = k_rate * dt * two_r_over_c;
                dcyc -= std::floor(dcyc);
                double wr = std::cos(kTwoPi * dcyc), wi = -std::sin(kTwoPi * dcyc);
                for (int m0 = 0; m0 < ns; m0 += kAnchor) {
                    double cyc = (f0 + k_rate * t[m0]) * two_r_over_c;
                    cyc -= std::floor(cyc);
                    double zr = std::cos(kTwoPi * cyc), zi = -std::sin(kTwoPi * cyc);
                    // fold the amplitude into the phasor
                    double pr = ar * zr - ai * zi, pi = ar * zi + ai * zr;
                    int m1 = std::min(ns, m0 + kAnchor);
                    for (int m = m0; m < m1; ++m) {
                        row[2 * m] += pr;
                        row[2 * m + 1] += pi;
                        double nr = pr * wr - pi * wi;
                        pi = pr * wi + pi * wr;
                        pr = nr;
                    }
                }
            } else {
                for (int m = 0; m < ns; ++m) {
                    double cyc = (f0 + k_rate * t[m]) * two_r_over_c;
                    cyc -= std::floor(cyc);
                    double zr = std::cos(kTwoPi * cyc), zi = -std::sin(kTwoPi * cyc);
                    row[2 * m] += ar * zr - ai * zi;
                    row[2 * m + 1] += ar * zi + ai * zr;
                }
            }
        }
    }
}

} // namespace

extern "C" {

// Beat signal of nk point scatterers for a stop-and-hop track.
//   track        [nc][3]  phase-centre position per chirp, m
//   sample_times [ns]     s, relative to chirp start
//   scat         [nk][3]  scatterer positions, m
//   amp          [nk][2]  complex reflectivities
//   out          [nc][ns][2] complex128 result (overwritten)
// Returns 0 on success, -1 on bad arguments.
int tdcsim_beat(const double *track, int nc, int ns, const double *sample_times, double f0,
                double k_rate, double c, const double *scat, const double *amp, int nk, double *out,
                int nthreads)
{
    if (!track || !sample_times || !out || nc <= 0 || ns <= 0 || nk < 0 || c <= 0.0) return -1;
    if (nk > 0 && (!scat || !amp)) return -1;
    double dt = 0.0;
    bool uniform = uniform_times(sample_times, ns, &dt);
    int nt = std::max(1, std::min(nthreads > 0 ? nthreads : 1, nc));
    std::vector<std::thread> pool;
    int per = (nc + nt - 1) / nt;
    for (int i = 0; i < nt; ++i) {
        int n0 = i * per, n1 = std::min(nc, n0 + per);
        if (n0 >= n1) break;
        pool.emplace_back(sim_rows, track, n0, n1, ns, sample_times, f0, k_rate, c, scat, amp, nk,
                          out, uniform, dt);
    }
    for (auto &th : pool) th.join();
    return 0;
}

} // extern "C"
