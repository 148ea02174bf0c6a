// The full variant's reverse mode (blochparty_amd/csrc/full_device.h, compiled here for the host) against
// Richardson-extrapolated central differences of its own forward, over random sequences of RF (some felt by the bound
// pool only) and free segments (some long enough for the exchange's closed-form branch), for every combination of a
// bound pool or none, no / leading / exact concomitant fields, and the Lorentzian (narrow and broad) or a tabulated
// super-Lorentzian lineshape. L = c . (state after the last segment). Prints one line per configuration and the worst
// relative error; exits 1 when it is above the bound given as argv[1] (default 1e-4).
//
// Errors: a spin parameter's relative to its own difference; a segment parameter's relative to the largest difference
// of its kind over the sequence. Differences below the rounding floor (|dL/dv| |v| < 1e-7 for spin parameters, a kind
// whose largest |dL/dv| is under 1e-8) are not judged.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../blochparty_amd/csrc/full_device.h"

using namespace hb::full;

namespace {

std::vector<float> table;

struct Case {
    Model md;
    Params p;
    std::vector<Seg> segs;
    St m0, c;
};

double run(const Case& k) {
    St m = k.m0;
    for (const Seg& q : k.segs) m = step(k.md, k.p, q, m).after;
    return k.c.x * m.x + k.c.y * m.y + k.c.z * m.z + k.c.b * m.b;
}

template <class F>
double richardson(double& v, double h, F f) {
    const double v0 = v;
    auto D = [&](double e) {
        v = v0 + e;
        const double a = f();
        v = v0 - e;
        const double b = f();
        v = v0;
        return (a - b) / (2 * e);
    };
    return (4 * D(h / 2) - D(h)) / 3;
}

void moments(Seg& q) { q.mx = q.gx * q.dt, q.my = q.gy * q.dt, q.mz = q.gz * q.dt, q.b0a = q.b0 * q.dt; }

}  // namespace

int main(int argc, char** argv) {
    const double bound = argc > 1 ? atof(argv[1]) : 1e-4;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1, 1);
    for (int i = 0; i < 200; ++i) table.push_back((float)(-0.5 - 0.02 * i - 0.1 * sin(0.3 * i)));
    double worst = 0.0;
    int fails = 0;
    for (int cfg = 0; cfg < 12; ++cfg) {
        Case k;
        const int line = cfg % 3;  // 0 narrow Lorentzian, 1 broad Lorentzian, 2 super-Lorentzian table
        k.md.pool = (cfg / 3) % 2;
        k.md.conc = (cfg / 6 + cfg) % 3;
        k.md.line = {table.data(), log(1e-3), 0.05, line == 2 ? 20.0 : 0.0, 200, line == 2 ? 2 : 1};
        Params& p = k.p;
        p.px = 0.05 * U(rng), p.py = 0.05 * U(rng), p.pz = 0.01 * U(rng);
        p.pd = 0.8 + 0.1 * U(rng), p.r1 = 1.0 + 0.3 * U(rng), p.r2 = 12 + 3 * U(rng), p.db0 = 30 * U(rng);
        p.b1 = 1.0 + 0.1 * U(rng);
        p.f = 0.12 + 0.05 * U(rng), p.k = 25 + 5 * U(rng), p.r1b = 1.0 + 0.2 * U(rng);
        p.t2b = line == 1 ? 2e-4 * (1 + 0.2 * U(rng)) : 1.1e-5 + 2e-6 * U(rng);
        p.dw = -150 + 20 * U(rng);
        for (double& c : p.cc) c = 0.0;
        if (k.md.conc == kLeading)
            for (int i = 0; i < 6; ++i) p.cc[i] = 1e-6 * (1 + U(rng));
        if (k.md.conc == kExact) {
            for (int i = 0; i < 6; ++i) p.cc[i] = 0.05 * U(rng);
            p.cc[6] = 0.35 * 42.577e6;
        }
        for (int s = 0; s < 40; ++s) {
            Seg q{};
            const bool rf = s % 4 != 3;
            q.flags = rf ? kRf : 0;
            if (rf && s % 7 == 2) q.flags |= kBoundOnly;
            q.dt = rf ? 2e-5 * (1 + 0.5 * U(rng)) : (s % 8 == 7 ? 0.08 : 1e-3) * (1 + 0.5 * U(rng));
            q.rfx = rf ? 400 * U(rng) : 0, q.rfy = rf ? 400 * U(rng) : 0;
            q.gx = 2e3 * U(rng), q.gy = 2e3 * U(rng), q.gz = 4e3 * U(rng), q.b0 = 10 * U(rng);
            q.frf = rf ? 500 * U(rng) : 0;
            moments(q);
            k.segs.push_back(q);
        }
        k.m0 = {0.1 * U(rng), 0.1 * U(rng), p.pd * 0.9, p.f * p.pd * 0.8};
        k.c = {U(rng), U(rng), U(rng), k.md.pool ? U(rng) : 0.0};
        // the reverse sweep
        std::vector<St> before;
        St m = k.m0;
        for (const Seg& q : k.segs) {
            before.push_back(m);
            m = step(k.md, p, q, m).after;
        }
        St l = k.c;
        double g[N_G] = {0};
        std::vector<std::array<double, N_S>> sg(k.segs.size());
        for (int s = (int)k.segs.size() - 1; s >= 0; --s) {
            double t[N_S] = {0};
            const Step st = step(k.md, p, k.segs[s], before[s]);
            step_reverse(k.md, p, k.segs[s], before[s], st, l, g, t);
            for (int i = 0; i < N_S; ++i) sg[s][i] = t[i];
        }
        double err = 0.0;
        // spin parameters
        double* ptr[N_G] = {&p.px,    &p.py,    &p.pz,    &p.db0,   &p.pd,    &p.r1, &p.r2, &p.b1,  &p.cc[0], &p.cc[1],
                            &p.cc[2], &p.cc[3], &p.cc[4], &p.cc[5], &p.cc[6], &p.f,  &p.k,  &p.r1b, &p.t2b,   &p.dw};
        for (int i = 0; i < N_G; ++i) {
            const double v = *ptr[i], scale = fabs(v) > 1e-12 ? fabs(v) : 1.0;
            const double fd = richardson(*ptr[i], 1e-4 * scale, [&] { return run(k); });
            if (fabs(fd) * scale > 1e-7) err = std::max(err, fabs(fd - g[i]) / fabs(fd));
        }
        // the initial state
        double* mp[4] = {&k.m0.x, &k.m0.y, &k.m0.z, &k.m0.b};
        const double lv[4] = {l.x, l.y, l.z, l.b};
        for (int i = 0; i < 4; ++i) {
            const double fd = richardson(*mp[i], 1e-5, [&] { return run(k); });
            if (fabs(fd) > 1e-9) err = std::max(err, fabs(fd - lv[i]) / fabs(fd));
        }
        // segment parameters
        double smax[N_S] = {0}, serr[N_S] = {0};
        for (size_t s = 0; s < k.segs.size(); ++s) {
            Seg& q = k.segs[s];
            double* sp[N_S] = {&q.dt, &q.rfx, &q.rfy, &q.gx, &q.gy, &q.gz, &q.b0, &q.frf};
            for (int i = 0; i < N_S; ++i) {
                const double v = *sp[i], scale = fabs(v) > 1e-12 ? fabs(v) : 1.0;
                const double fd = richardson(*sp[i], 1e-4 * scale, [&] {
                    moments(q);
                    return run(k);
                });
                moments(q);
                smax[i] = std::max(smax[i], fabs(fd));
                serr[i] = std::max(serr[i], fabs(fd - sg[s][i]));
            }
        }
        for (int i = 0; i < N_S; ++i)
            if (smax[i] > 1e-8) err = std::max(err, serr[i] / smax[i]);
        printf("pool %d, concomitant %d, lineshape %s: worst relative error %.2e\n", (int)k.md.pool, k.md.conc,
               line == 0 ? "Lorentzian (narrow)" : line == 1 ? "Lorentzian (broad)" : "super-Lorentzian table", err);
        worst = std::max(worst, err);
        fails += err > bound;
    }
    printf("worst %.3e\n", worst);
    return fails ? 1 : 0;
}
