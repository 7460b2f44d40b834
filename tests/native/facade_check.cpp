// Drives the C++ facades exactly as the reference's electric.cpp:294-392
// drives bbfmm: sixteen bbfmm::kernel objects initialised with the weighted
// charges and run one after another, their results combined into the 2N
// output, and GMRES(std::function) on that operator (gmres.cpp:50-164).
// Compares them with the plan's own matvec (ibim_matvec) and device GMRES
// (GMRES(plan, ...)) on a seeded surface-like point set; prints one line per
// comparison and exits non-zero if a bar is missed.
//   facade_check [n] [seed]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "bbfmm.h"
#include "gmres.h"

using bbfmm::kernel;
using bbfmm::point;
using bbfmm::Vector;

namespace {

double rel(const Vector &a, const Vector &b) {
    double d = 0, n = 0;
    for (int i = 0; i < a.row(); ++i) {
        d += (a(i) - b(i)) * (a(i) - b(i));
        n += b(i) * b(i);
    }
    return std::sqrt(d / n);
}

struct Case {
    std::vector<point> pts;
    std::vector<double> weight, normalX, normalY, normalZ;
};

// points on a shell of radius 2 (thickness 0.2) with unit radial normals
Case shell(int n, unsigned seed) {
    std::mt19937_64 g(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    Case c;
    for (int i = 0; i < n; ++i) {
        double v[3] = {nd(g), nd(g), nd(g)};
        const double r = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (double &x : v) x /= r;
        const double rad = 2.0 + 0.2 * (ud(g) - 0.5);
        c.pts.emplace_back(rad * v[0], rad * v[1], rad * v[2]);
        c.normalX.push_back(v[0]);
        c.normalY.push_back(v[1]);
        c.normalZ.push_back(v[2]);
        c.weight.push_back(0.01 + 0.01 * ud(g));
    }
    return c;
}

}  // namespace

int main(int argc, char **argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 2000;
    const unsigned seed = argc > 2 ? (unsigned)std::atoi(argv[2]) : 5;
    const int np = 4, maxPoint = 32, maxLevel = 10;
    const double kappa = 0.1257, dI = 2.0, dE = 80.0, vacant = 0.05;
    Case c = shell(N, seed);
    std::vector<point> &source = c.pts, &target = c.pts;
    std::vector<double> &weight = c.weight, &normalX = c.normalX, &normalY = c.normalY, &normalZ = c.normalZ;

    // the operator as electric.cpp:294-388 forms it, kernel by kernel
    auto dmapping = [&](Vector &_phi) {
        kernel _K1x, _K1y, _K1z, _K2, _K3xx, _K3yy, _K3zz, _K3xy, _K3yz, _K3zx, _K3yx, _K3xz, _K3zy, _K4x, _K4y, _K4z;
        kernel *all[16] = {&_K1x, &_K1y, &_K1z, &_K2, &_K3xx, &_K3yy, &_K3zz, &_K3xy,
                           &_K3yz, &_K3zx, &_K3yx, &_K3zy, &_K3xz, &_K4x, &_K4y, &_K4z};
        const int ids[16] = {bbfmm::K1x, bbfmm::K1y, bbfmm::K1z, bbfmm::K2, bbfmm::K3xx, bbfmm::K3yy,
                             bbfmm::K3zz, bbfmm::K3xy, bbfmm::K3yz, bbfmm::K3zx, bbfmm::K3yx, bbfmm::K3zy,
                             bbfmm::K3xz, bbfmm::K4x, bbfmm::K4y, bbfmm::K4z};
        for (int k = 0; k < 16; ++k) {  // the kernel closures of electric.cpp:67-291, by id
            all[k]->kernelId = ids[k];
            all[k]->params.kappa = kappa;
            all[k]->params.eps_in = dI;
            all[k]->params.eps_out = dE;
            all[k]->params.vacant_radius = vacant;
        }
        Vector pphi_pn(N), phiX(N), phiY(N), phiZ(N);
        for (int id = 0; id < N; ++id) {
            pphi_pn(id) = _phi(id + N) * weight[id];
            phiX(id) = _phi(id) * normalX[id] * weight[id];
            phiY(id) = _phi(id) * normalY[id] * weight[id];
            phiZ(id) = _phi(id) * normalZ[id] * weight[id];
        }
        _K1x.initialize(np, source, target, phiX, N, N, maxPoint, maxLevel);
        _K1y.initialize(np, source, target, phiY, N, N, maxPoint, maxLevel);
        _K1z.initialize(np, source, target, phiZ, N, N, maxPoint, maxLevel);
        _K2.initialize(np, source, target, pphi_pn, N, N, maxPoint, maxLevel);
        _K3xx.initialize(np, source, target, phiX, N, N, maxPoint, maxLevel);
        _K3yy.initialize(np, source, target, phiY, N, N, maxPoint, maxLevel);
        _K3zz.initialize(np, source, target, phiZ, N, N, maxPoint, maxLevel);
        _K3xy.initialize(np, source, target, phiX, N, N, maxPoint, maxLevel);
        _K3yz.initialize(np, source, target, phiY, N, N, maxPoint, maxLevel);
        _K3zx.initialize(np, source, target, phiZ, N, N, maxPoint, maxLevel);
        _K3yx.initialize(np, source, target, phiY, N, N, maxPoint, maxLevel);
        _K3zy.initialize(np, source, target, phiZ, N, N, maxPoint, maxLevel);
        _K3xz.initialize(np, source, target, phiX, N, N, maxPoint, maxLevel);
        _K4x.initialize(np, source, target, pphi_pn, N, N, maxPoint, maxLevel);
        _K4y.initialize(np, source, target, pphi_pn, N, N, maxPoint, maxLevel);
        _K4z.initialize(np, source, target, pphi_pn, N, N, maxPoint, maxLevel);
        Vector r1x, r1y, r1z, r2, r3xx, r3yy, r3zz, r3xy, r3yz, r3zx, r3yx, r3zy, r3xz, r4x, r4y, r4z;
        _K1x.run(r1x);
        _K1y.run(r1y);
        _K1z.run(r1z);
        _K2.run(r2);
        _K3xx.run(r3xx);
        _K3yy.run(r3yy);
        _K3zz.run(r3zz);
        _K3xy.run(r3xy);
        _K3yx.run(r3yx);
        _K3xz.run(r3xz);
        _K3zx.run(r3zx);
        _K3yz.run(r3yz);
        _K3zy.run(r3zy);
        _K4x.run(r4x);
        _K4y.run(r4y);
        _K4z.run(r4z);
        Vector output(2 * N);
        for (int id = 0; id < N; ++id) {
            const double t1 = normalX[id] * (r3xx(id) + r3yx(id) + r3zx(id)) +
                              normalY[id] * (r3yy(id) + r3xy(id) + r3zy(id)) +
                              normalZ[id] * (r3zz(id) + r3xz(id) + r3yz(id));
            const double t2 = normalX[id] * r4x(id) + normalY[id] * r4y(id) + normalZ[id] * r4z(id);
            output(id) = _phi(id) + (r1x(id) + r1y(id) + r1z(id) - r2(id)) / (0.5 * (1 + dE / dI));
            output(id + N) = _phi(id + N) + (t1 - t2) / (0.5 * (1 + dI / dE));
        }
        return output;
    };
    auto FullMap = [&](Vector &phi) { return dmapping(phi); };

    // the same operator as one plan
    std::vector<double> xyz(3 * (size_t)N), nrm(3 * (size_t)N);
    for (int i = 0; i < N; ++i) {
        xyz[3 * i] = source[i].x;
        xyz[3 * i + 1] = source[i].y;
        xyz[3 * i + 2] = source[i].z;
        nrm[3 * i] = normalX[i];
        nrm[3 * i + 1] = normalY[i];
        nrm[3 * i + 2] = normalZ[i];
    }
    ibim_params prm{};
    prm.p = np;
    prm.max_points = maxPoint;
    prm.max_level = maxLevel;
    prm.kappa = kappa;
    prm.eps_in = dI;
    prm.eps_out = dE;
    prm.vacant_radius = vacant;
    prm.precision = IBIM_PREC_FP64;
    prm.kernel_mode = IBIM_KERNEL_PB;
    ibim_plan *plan = nullptr;
    ibim::check(ibim_plan_create(&prm, N, xyz.data(), nrm.data(), weight.data(), &plan), "ibim_plan_create");

    int fails = 0;
    Vector x(2 * N);
    std::mt19937_64 g(seed + 1);
    std::uniform_real_distribution<double> ud(-1.0, 1.0);
    for (int i = 0; i < 2 * N; ++i) x(i) = ud(g);
    Vector y_facade = FullMap(x), y_plan(2 * N);
    ibim::check(ibim_matvec(plan, x.data(), y_plan.data()), "ibim_matvec");
    const double e_mv = rel(y_facade, y_plan);
    std::printf("facade dmapping vs ibim_matvec: rel %.3e\n", e_mv);
    fails += !(e_mv < 1e-10);

    // GMRES(std::function) on the facade operator vs GMRES(plan)
    Vector b(2 * N);
    for (int i = 0; i < 2 * N; ++i) b(i) = ud(g);
    Vector xf(2 * N), xp(2 * N);
    const int rf = GMRES(FullMap, xf, b, 10, 200, 1e-10);
    ibim_solve_info info{};
    const int rp = GMRES(plan, xp, b, 10, 200, 1e-10, &info);
    const double e_sol = rel(xf, xp);
    std::printf("GMRES(std::function) vs GMRES(plan): rc %d / %d, solution rel %.3e, plan iterations %d\n", rf, rp,
                e_sol, info.iterations);
    fails += !(rf == 0 && rp == 0 && e_sol < 1e-8);

    // a closure-only kernel fails loudly instead of falling back to the CPU
    bool threw = false;
    try {
        kernel k;
        k.eval = [](point &, point &) { return 1.0; };
        Vector q(N);
        k.initialize(np, source, target, q, N, N, maxPoint, maxLevel);
    } catch (const ibim::Error &) {
        threw = true;
    }
    std::printf("closure-only kernel rejected: %s\n", threw ? "yes" : "no");
    fails += !threw;
    ibim_plan_destroy(plan);
    std::printf("%s\n", fails ? "FACADE CHECK FAILED" : "FACADE CHECK OK");
    return fails ? 1 : 0;
}
