// Host checks of the ingest helpers in libsarlabtdc_amd.so (no device calls):
// tdc::parallel_for visits every index exactly once, also when called from
// several threads at once, and tdc::scan_antenna agrees with a plain serial
// pass over the same per-sample antenna states. Prints "OK" and exits 0.
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
