// Host checks of the ingest helpers in libsarlabtdc_amd.so (no device calls):
// tdc::parallel_for visits every index exactly once, also when called from
// several threads at once, and tdc::scan_antenna agrees with a plain serial
// pass over the same per-sample antenna states. Prints "OK" and exits 0.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "tdc_internal.h"

namespace {

int failures = 0;

void expect(bool ok, const char *what)
{
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++failures;
    }
}

bool visits_once(size_t n, size_t grain)
{
    std::vector<std::atomic<int>> seen(n);
    for (auto &v : seen) v.store(0);
    std::atomic<unsigned long long> sum{0};
    tdc::parallel_for(n, grain, [&](size_t b, size_t e) {
        unsigned long long s = 0;
        for (size_t i = b; i < e; ++i) {
            seen[i].fetch_add(1);
            s += i;
        }
        sum.fetch_add(s);
    });
    for (auto &v : seen)
        if (v.load() != 1) return false;
    return n == 0 || sum.load() == static_cast<unsigned long long>(n) * (n - 1) / 2;
}

// The same pass written out serially.
tdc::AntennaScan serial_scan(const float *pos, const float *vel, const float *att, int npri, int ns)
{
    tdc::AntennaScan s;
    s.has_vel = vel != nullptr;
    s.has_att = att != nullptr;
    for (int a = 0; a < 3; ++a) s.lo[a] = s.hi[a] = pos[a];
    for (int n = 0; n < npri; ++n) {
        for (int m = 0; m < ns; ++m) {
            const size_t e = static_cast<size_t>(n) * ns + m, e0 = static_cast<size_t>(n) * ns;
            for (int a = 0; a < 3; ++a) {
                const float v = pos[3 * e + a];
                if (!std::isfinite(v)) s.finite = false;
                if (v < s.lo[a]) s.lo[a] = v;
                if (v > s.hi[a]) s.hi[a] = v;
            }
            if (std::memcmp(pos + 3 * e, pos + 3 * e0, 12) != 0) s.pos_const = false;
            const double ox = double(pos[3 * e]) - double(pos[3 * e0]),
                         oy = double(pos[3 * e + 1]) - double(pos[3 * e0 + 1]),
                         oz = double(pos[3 * e + 2]) - double(pos[3 * e0 + 2]);
            const double d2 = ox * ox + oy * oy + oz * oz;
            if (d2 > s.dmax2) s.dmax2 = d2;
            if (vel) {
                double v2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    if (!std::isfinite(vel[3 * e + a])) s.finite = false;
                }
                v2 = double(vel[3 * e]) * vel[3 * e] + double(vel[3 * e + 1]) * vel[3 * e + 1] +
                     double(vel[3 * e + 2]) * vel[3 * e + 2];
                if (v2 > s.vmax2) s.vmax2 = v2;
                if (std::memcmp(vel + 3 * e, vel + 3 * e0, 12) != 0) s.vel_const = false;
            }
            if (att) {
                double q2 = 0.0;
                for (int a = 0; a < 4; ++a) {
                    if (!std::isfinite(att[4 * e + a])) s.finite = false;
                    q2 += double(att[4 * e + a]) * att[4 * e + a];
                }
                if (q2 > s.qmax2) s.qmax2 = q2;
                if (std::memcmp(att + 4 * e, att + 4 * e0, 16) != 0) s.att_const = false;
            }
        }
    }
    return s;
}

bool same(const tdc::AntennaScan &a, const tdc::AntennaScan &b)
{
    return a.has_vel == b.has_vel && a.has_att == b.has_att && a.finite == b.finite &&
           a.pos_const == b.pos_const && a.vel_const == b.vel_const &&
           a.att_const == b.att_const && a.vmax2 == b.vmax2 && a.qmax2 == b.qmax2 &&
           a.dmax2 == b.dmax2 &&
           std::memcmp(a.lo, b.lo, sizeof a.lo) == 0 && std::memcmp(a.hi, b.hi, sizeof a.hi) == 0;
}

struct Antenna {
    int npri, ns;
    std::vector<float> pos, vel, att;
};

// Stop-and-hop states: one value per chirp, repeated over its samples.
Antenna make_antenna(int npri, int ns, std::mt19937 &rng)
{
    std::normal_distribution<float> nd;
    Antenna a{npri, ns, {}, {}, {}};
    a.pos.resize(3 * static_cast<size_t>(npri) * ns);
    a.vel.resize(a.pos.size());
    a.att.resize(4 * static_cast<size_t>(npri) * ns);
    for (int n = 0; n < npri; ++n) {
        const float p[3] = {nd(rng), nd(rng), nd(rng)}, v[3] = {nd(rng), nd(rng), nd(rng)};
        const float q[4] = {1.0F + 0.01F * nd(rng), 0.01F * nd(rng), 0.01F * nd(rng), 0.01F * nd(rng)};
        for (int m = 0; m < ns; ++m) {
            const size_t e = static_cast<size_t>(n) * ns + m;
            std::memcpy(&a.pos[3 * e], p, 12);
            std::memcpy(&a.vel[3 * e], v, 12);
            std::memcpy(&a.att[4 * e], q, 16);
        }
    }
    return a;
}

void check_scan(const Antenna &a, bool with_vel, bool with_att, const char *what)
{
    const float *v = with_vel ? a.vel.data() : nullptr, *q = with_att ? a.att.data() : nullptr;
    expect(same(tdc::scan_antenna(a.pos.data(), v, q, a.npri, a.ns),
                serial_scan(a.pos.data(), v, q, a.npri, a.ns)),
           what);
}

} // namespace

// Origin-relative chirp positions of an arc of radius `radius` about the
// origin, npri chirps a `step` apart, with Gaussian jitter.
std::vector<tdc::ChirpPos> arc_track(int npri, double radius, double step, double jitter,
                                     std::mt19937 &rng)
{
    std::normal_distribution<double> n01(0.0, 1.0);
    std::vector<tdc::ChirpPos> c(npri);
    for (int n = 0; n < npri; ++n) {
        const double ang = (n - npri / 2) * step / radius;
        const double x = radius * std::sin(ang) + jitter * n01(rng);
        const double y = -radius * std::cos(ang) + jitter * n01(rng);
        const double z = 0.3 + jitter * n01(rng);
        c[n] = tdc::ChirpPos{x, y, z, x * x + y * y + z * z};
    }
    return c;
}

// The range increments the range-profile tile forms in FP32 (the same
// operations) against the FP64 ranges, over blocks of 8 chirps: the largest
// step error in metres.
double incremental_error(const std::vector<tdc::ChirpPos> &c, const std::vector<double> &px)
{
    double worst = 0.0;
    const int npri = static_cast<int>(c.size());
    for (size_t p = 0; p + 2 < px.size(); p += 3) {
        const float qx = static_cast<float>(-px[p]), qy = static_cast<float>(-px[p + 1]),
                    qz = static_cast<float>(-px[p + 2]);
        auto range = [&](int n) {
            const double dx = c[n].x - px[p], dy = c[n].y - px[p + 1], dz = c[n].z - px[p + 2];
            return std::sqrt(dx * dx + dy * dy + dz * dz);
        };
        for (int nb = 0; nb < npri; nb += 8) {
            float s = static_cast<float>(range(nb));
            float is = 1.0F / s;
            for (int n = nb + 1; n < std::min(npri, nb + 8); ++n) {
                const float dx = static_cast<float>(c[n].x - c[n - 1].x);
                const float dy = static_cast<float>(c[n].y - c[n - 1].y);
                const float dz = static_cast<float>(c[n].z - c[n - 1].z);
                const float dw = static_cast<float>(0.5 * (c[n].pad - c[n - 1].pad));
                const float hdelta = std::fmaf(qz, dz, std::fmaf(qy, dy, std::fmaf(qx, dx, dw)));
                const float e0 = hdelta * is;
                const float u = e0 * is;
                const float dr = e0 * std::fmaf(u, std::fmaf(0.5F, u, -0.5F), 1.0F);
                worst = std::max(worst, std::fabs(static_cast<double>(dr) - (range(n) - range(n - 1))));
                s += dr;
                is = 1.0F / s;
            }
        }
    }
    return worst;
}

// The largest wavenumber for which profile_incremental_ok holds (bisection):
// the bound it applies per step is then kIncrementalCycles / kmax metres.
double kmax_limit(const std::vector<tdc::ChirpPos> &c, const double gmin[3], const double gmax[3])
{
    double lo = 0.0, hi = 1e12;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (tdc::profile_incremental_ok(c.data(), static_cast<int>(c.size()), gmin, gmax, mid)) lo = mid;
        else hi = mid;
    }
    return lo;
}

void check_incremental(std::mt19937 &rng)
{
    // a 0.5 m grid at the origin, a 1.5 m arc, 1 mm steps with 0.5 mm jitter
    std::vector<tdc::ChirpPos> c = arc_track(1024, 1.5, 1e-3, 5e-4, rng);
    std::uniform_real_distribution<double> u(-0.25, 0.25);
    std::vector<double> px(3 * 400);
    for (double &v : px) v = u(rng);
    const double gmin[3] = {-0.25, -0.25, -0.25}, gmax[3] = {0.25, 0.25, 0.25};
    expect(tdc::profile_incremental_ok(c.data(), 1024, gmin, gmax, 527.0),
           "incremental range allowed for a 1.5 m arc with 1 mm steps at 77 GHz");
    const double k = kmax_limit(c, gmin, gmax);
    const double err = incremental_error(c, px);
    std::printf("incremental range: largest step error %.3e m, bound %.3e m\n", err,
                tdc::kIncrementalCycles / k);
    expect(err <= tdc::kIncrementalCycles / k, "FP32 range steps within the bound");
    // the same with 2 cm steps: the series' remainder fails the bound
    std::vector<tdc::ChirpPos> big = arc_track(64, 1.5, 2e-2, 0.0, rng);
    expect(!tdc::profile_incremental_ok(big.data(), 64, gmin, gmax, 527.0),
           "2 cm steps take the FP64 range");
    // pixels reaching the antenna's path: no positive distance bound
    const double gmin2[3] = {-2.0, -2.0, -2.0}, gmax2[3] = {2.0, 2.0, 2.0};
    expect(!tdc::profile_incremental_ok(c.data(), 1024, gmin2, gmax2, 527.0),
           "overlapping boxes take the FP64 range");
    // an antenna 10 km away: delta grows with the range, its rounding error
    // over 2 R does not; a wavenumber 1000 times larger fails the bound
    std::vector<tdc::ChirpPos> far = c;
    for (auto &q : far) {
        q.y -= 1e4;
        q.pad = q.x * q.x + q.y * q.y + q.z * q.z;
    }
    expect(tdc::profile_incremental_ok(far.data(), 1024, gmin, gmax, 527.0),
           "a distant antenna keeps the incremental range");
    const double kf = kmax_limit(far, gmin, gmax);
    const double errf = incremental_error(far, px);
    std::printf("distant antenna: largest step error %.3e m, bound %.3e m\n", errf,
                tdc::kIncrementalCycles / kf);
    expect(errf <= tdc::kIncrementalCycles / kf,
           "FP32 range steps of a distant antenna within the bound");
    expect(!tdc::profile_incremental_ok(c.data(), 1024, gmin, gmax, 527e3),
           "a 1000 times larger wavenumber takes the FP64 range");
    expect(!tdc::profile_incremental_ok(c.data(), 1, gmin, gmax, 527.0), "one chirp: FP64");
    std::vector<tdc::ChirpPos> bad = c;
    bad[7].x = std::numeric_limits<double>::quiet_NaN();
    expect(!tdc::profile_incremental_ok(bad.data(), 1024, gmin, gmax, 527.0), "NaN position: FP64");
    const double gnan[3] = {std::numeric_limits<double>::quiet_NaN(), 0.0, 0.0};
    expect(!tdc::profile_incremental_ok(c.data(), 1024, gnan, gmax, 527.0), "NaN grid bound: FP64");

    // grid_bounds against a serial pass, and NaN
    std::vector<double> g(3 * 100003);
    for (double &v : g) v = u(rng) * 7.0;
    double lo[3], hi[3];
    tdc::grid_bounds(g.data(), g.size() / 3, lo, hi);
    bool same = true;
    for (int a = 0; a < 3; ++a) {
        double l = g[a], h = g[a];
        for (size_t i = 0; i < g.size() / 3; ++i) {
            l = std::min(l, g[3 * i + a]);
            h = std::max(h, g[3 * i + a]);
        }
        same = same && l == lo[a] && h == hi[a];
    }
    expect(same, "grid_bounds equals a serial pass");
    g[3 * 5000 + 1] = std::numeric_limits<double>::quiet_NaN();
    tdc::grid_bounds(g.data(), g.size() / 3, lo, hi);
    expect(std::isnan(lo[0]) && std::isnan(hi[2]), "grid_bounds reports NaN");
}

int main()
{
    for (size_t n : {size_t(0), size_t(1), size_t(5), size_t(1000), size_t(1000003)})
        for (size_t grain : {size_t(1), size_t(7), size_t(4096)})
            expect(visits_once(n, grain), "parallel_for visits every index once");

    // calls from several threads at once are serialised, none is lost
    std::vector<std::thread> callers;
    std::atomic<int> bad{0};
    for (int t = 0; t < 4; ++t)
        callers.emplace_back([&] {
            for (int k = 0; k < 200; ++k)
                if (!visits_once(20000 + 37 * k, 64)) bad.fetch_add(1);
        });
    for (auto &c : callers) c.join();
    expect(bad.load() == 0, "parallel_for from four threads at once");

    std::mt19937 rng(11);
    check_incremental(rng);
    Antenna a = make_antenna(1021, 64, rng);
    check_scan(a, true, true, "stop-and-hop states");
    expect(tdc::scan_antenna(a.pos.data(), a.vel.data(), a.att.data(), a.npri, a.ns).pos_const,
           "stop-and-hop positions are constant per chirp");
    check_scan(a, false, false, "positions only");
    check_scan(a, true, false, "no attitude");

    // one sample of one chirp moves: only that array stops being constant
    Antenna b = a;
    b.pos[3 * (static_cast<size_t>(700) * b.ns + 13) + 1] += 1e-3F;
    check_scan(b, true, true, "one position differs");
    tdc::AntennaScan sb = tdc::scan_antenna(b.pos.data(), b.vel.data(), b.att.data(), b.npri, b.ns);
    expect(!sb.pos_const && sb.vel_const && sb.att_const, "only the positions vary");
    Antenna c = a;
    c.vel[3 * (static_cast<size_t>(1020) * c.ns + 63)] *= 2.0F;
    c.att[4 * (static_cast<size_t>(3) * c.ns + 1) + 2] += 1e-4F;
    check_scan(c, true, true, "velocity and attitude differ");
    tdc::AntennaScan sc = tdc::scan_antenna(c.pos.data(), c.vel.data(), c.att.data(), c.npri, c.ns);
    expect(sc.pos_const && !sc.vel_const && !sc.att_const, "velocity and attitude vary");

    // +0.0 and -0.0 differ bit for bit
    Antenna d = a;
    for (int m = 0; m < d.ns; ++m) d.pos[3 * (static_cast<size_t>(5) * d.ns + m)] = 0.0F;
    d.pos[3 * (static_cast<size_t>(5) * d.ns + 9)] = -0.0F;
    check_scan(d, true, true, "signed zero");
    expect(!tdc::scan_antenna(d.pos.data(), nullptr, nullptr, d.npri, d.ns).pos_const,
           "signed zeros are different bits");

    // a NaN anywhere clears the finite flag
    Antenna e = a;
    e.att[4 * (static_cast<size_t>(512) * e.ns + 30) + 3] = std::numeric_limits<float>::quiet_NaN();
    check_scan(e, true, true, "NaN attitude");
    expect(!tdc::scan_antenna(e.pos.data(), e.vel.data(), e.att.data(), e.npri, e.ns).finite,
           "NaN attitude is not finite");
    expect(tdc::scan_antenna(e.pos.data(), e.vel.data(), nullptr, e.npri, e.ns).finite,
           "without attitude the states are finite");

    if (failures) return 1;
    std::printf("OK\n");
    return 0;
}
