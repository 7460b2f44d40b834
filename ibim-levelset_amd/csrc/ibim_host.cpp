// ibim_host.cpp -- host-side molecule helpers of the C ABI.
//
// ibim_synthetic_globule: seeded stand-in molecules for the benchmark
// configurations.  A self-avoiding random walk of 3.8 A steps (the C-alpha
// spacing of a protein chain) inside a confinement sphere; a step is rejected
// if it leaves the sphere or comes within 3.0 A of an atom other than the one
// it starts from; after 200 rejected trials the walk continues from a random
// earlier atom.  Radii are U[1.2, 2.0] A and charges U[-0.5, 0.5] e, shifted
// uniformly so that the net charge is -2 e.  Random doubles are drawn as
// (u >> 11) * 2^-53 from std::mt19937_64, so the molecules are identical on
// every platform.
#include <algorithm>
#include <cmath>
#include <string>
#include <cstdint>
#include <random>
#include <unordered_map>
#include <vector>

#include "../../include/ibim.h"

namespace ibim {
void set_error(const std::string &msg);
}

namespace {

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    double u() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
};

struct CellHash {
    double cell;
    std::unordered_map<int64_t, std::vector<int>> cells;
    static int64_t key(int64_t a, int64_t b, int64_t c) {
        return ((a + (1 << 20)) << 42) | ((b + (1 << 20)) << 21) | (c + (1 << 20));
    }
    void insert(const double *p, int id) {
        cells[key((int64_t)std::floor(p[0] / cell), (int64_t)std::floor(p[1] / cell), (int64_t)std::floor(p[2] / cell))]
            .push_back(id);
    }
};

}  // namespace

extern "C" int ibim_synthetic_globule(int64_t m, uint64_t seed, double confine_radius, double *xyz, double *radii,
                                      double *charges) {
    if (m < 1 || !xyz || !radii || !charges || !(confine_radius > 0)) {
        ibim::set_error("bad argument to ibim_synthetic_globule");
        return IBIM_ERR_ARG;
    }
    const double step = 3.8, excl = 3.0, excl2 = excl * excl, R2 = confine_radius * confine_radius;
    Rng rng(seed);
    CellHash hash{excl, {}};
    std::vector<double> pos;
    pos.reserve(3 * (size_t)m);
    pos.insert(pos.end(), {0.0, 0.0, 0.0});
    hash.insert(pos.data(), 0);
    int64_t cur = 0;
    int64_t guard = 0;
    while ((int64_t)(pos.size() / 3) < m) {
        bool placed = false;
        for (int trial = 0; trial < 200 && !placed; ++trial) {
            const double z = 2.0 * rng.u() - 1.0, ph = 2.0 * M_PI * rng.u();
            const double s = std::sqrt(std::max(0.0, 1.0 - z * z));
            const double c[3] = {pos[3 * cur] + step * s * std::cos(ph), pos[3 * cur + 1] + step * s * std::sin(ph),
                                 pos[3 * cur + 2] + step * z};
            if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > R2) continue;
            const int64_t ca = (int64_t)std::floor(c[0] / excl), cb = (int64_t)std::floor(c[1] / excl),
                          cc = (int64_t)std::floor(c[2] / excl);
            bool ok = true;
            for (int64_t a = ca - 1; a <= ca + 1 && ok; ++a)
                for (int64_t b = cb - 1; b <= cb + 1 && ok; ++b)
                    for (int64_t q = cc - 1; q <= cc + 1 && ok; ++q) {
                        auto it = hash.cells.find(CellHash::key(a, b, q));
                        if (it == hash.cells.end()) continue;
                        for (int id : it->second) {
                            if (id == cur) continue;
                            const double dx = pos[3 * id] - c[0], dy = pos[3 * id + 1] - c[1], dz = pos[3 * id + 2] - c[2];
                            if (dx * dx + dy * dy + dz * dz < excl2) {
                                ok = false;
                                break;
                            }
                        }
                    }
            if (!ok) continue;
            const int id = (int)(pos.size() / 3);
            pos.insert(pos.end(), {c[0], c[1], c[2]});
            hash.insert(c, id);
            cur = id;
            placed = true;
        }
        if (!placed) {
            cur = (int64_t)(rng.u() * (double)(pos.size() / 3));
            if (++guard > 1000000 * (m + 10)) {
                ibim::set_error("synthetic globule: the walk jammed (confinement too small)");
                return IBIM_ERR_ARG;
            }
        }
    }
    double qsum = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        for (int q = 0; q < 3; ++q) xyz[3 * i + q] = pos[3 * i + q];
        radii[i] = 1.2 + 0.8 * rng.u();
    }
    for (int64_t i = 0; i < m; ++i) {
        charges[i] = rng.u() - 0.5;
        qsum += charges[i];
    }
    const double shift = (-2.0 - qsum) / (double)m;
    for (int64_t i = 0; i < m; ++i) charges[i] += shift;
    return IBIM_OK;
}
