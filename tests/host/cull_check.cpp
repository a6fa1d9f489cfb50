// Host check of the per-tile cull (the device code is host-callable): splat_may_reach, the region test of
// the transfer and element-adjoint culls, never rejects a region one of whose samples outside_support
// accepts, in FP64 or FP32, over random splats (rotating and static) and random sample clouds with
// emission times up to 1e4 M. Exit status 0 when no region was wrongly culled.
#include "../../kerrsplat_amd/csrc/ks_kernels.hpp"
#include <cstdio>
#include <random>
#include <vector>
using namespace ksd;
int main() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0, 1);
    std::normal_distribution<double> G(0, 1);
    long bad = 0, reach = 0, inside = 0, total = 0;
    for (int trial = 0; trial < 200000; ++trial) {
        SplatC c{}; 
        double r = 3 + 8 * U(rng), th = M_PI / 2 + 0.3 * G(rng), ph = 2 * M_PI * U(rng);
        c.x = r * sin(th) * cos(ph); c.y = r * sin(th) * sin(ph); c.z = r * cos(th);
        c.omega = (U(rng) < 0.1 ? 0.0 : (U(rng) < 0.5 ? -1 : 1) * (0.01 + 0.15 * U(rng)));
        c.t0 = 20 * G(rng);
        double smax = log(0.3 + 2 * U(rng));
        c.rsup2 = -2.0 * log(1e-6) * exp(2 * smax);
        SplatCf cf{}; cf.x = c.x; cf.y = c.y; cf.z = c.z; cf.omega = c.omega; cf.t0 = c.t0; cf.rsup2 = c.rsup2;
        // a region: samples scattered around a point near the splat's orbit, times over an interval
        double R0 = r + 4 * G(rng), ph0 = 2 * M_PI * U(rng), z0 = c.z + 3 * G(rng);
        double tbase = (U(rng) < 0.5 ? 1 : -1) * pow(10.0, 4 * U(rng)), tspan = 100 * U(rng);
        double spread = 0.1 + 3 * U(rng);
        int ns = 1 + (int)(64 * U(rng));
        std::vector<double> xs(ns), ys(ns), zs(ns), ts(ns);
        double lo[4] = {1e300, 1e300, 1e300, 1e300}, hi[4] = {-1e300, -1e300, -1e300, -1e300};
        for (int q = 0; q < ns; ++q) {
            xs[q] = R0 * cos(ph0) + spread * G(rng); ys[q] = R0 * sin(ph0) + spread * G(rng); zs[q] = z0 + spread * G(rng);
            ts[q] = tbase + tspan * U(rng);
            double v[4] = {xs[q], ys[q], zs[q], ts[q]};
            for (int k = 0; k < 4; ++k) { lo[k] = fmin(lo[k], v[k]); hi[k] = fmax(hi[k], v[k]); }
        }
        double px = 0.5 * (lo[0] + hi[0]), py = 0.5 * (lo[1] + hi[1]), pz = 0.5 * (lo[2] + hi[2]);
        double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        double R = 0.5 * sqrt(ex * ex + ey * ey + ez * ez);
        bool may64 = splat_may_reach(c, px, py, pz, R, lo[3], hi[3]);
        bool may32 = splat_may_reach(cf, px, py, pz, R, lo[3], hi[3]);
        bool in = false;
        for (int q = 0; q < ns; ++q) {
            double sp, cp; f32 sp32, cp32;
            if (!outside_support(c, ts[q], xs[q], ys[q], zs[q], sp, cp)) in = true;
            if (!outside_support(cf, f32(ts[q]), f32(xs[q]), f32(ys[q]), f32(zs[q]), sp32, cp32)) in = true;
        }
        ++total; inside += in; reach += may64;
        if (in && (!may64 || !may32)) { ++bad; if (bad < 5) printf("MISS trial %d: r %g omega %g t %g span %g\n", trial, r, c.omega, tbase, tspan); }
    }
    printf("regions %ld, with a sample inside %ld, kept by the cull %ld, wrongly culled %ld\n", total, inside, reach, bad);
    return bad != 0;
}
