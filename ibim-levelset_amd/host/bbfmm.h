// bbfmm.h -- the reference's FMM interface (bbfmm.h:24-270) over the GPU plan.
//
// bbfmm::kernel keeps initialize(...) / run(Vector&) with the same arguments
// and meaning.  The reference stores the kernel as a black-box closure
// (std::function<scalar_t(point&, point&)>); a host closure cannot run on the
// GPU, so a kernel names one of the 16 PB kernels of electric.cpp:67-291 by
// `kernelId` (with the physical constants in `params`).  A kernel with only a
// closure set throws instead of falling back to the CPU.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "ibim.hpp"

namespace bbfmm {

typedef double scalar_t;
typedef int index_t;

class point {
public:
    scalar_t x = 0, y = 0, z = 0;
    point() = default;
    point(scalar_t _x, scalar_t _y, scalar_t _z) : x(_x), y(_y), z(_z) {}
};

// column vector (linalg.h:11-132): zero-initialised, operator() access
class Vector {
public:
    explicit Vector(int row = 0) : d_((size_t)(row > 0 ? row : 0), 0.0) {}
    int row() const { return (int)d_.size(); }
    scalar_t &operator()(int i) { return d_[(size_t)i]; }
    scalar_t operator()(int i) const { return d_[(size_t)i]; }
    scalar_t *data() { return d_.data(); }
    const scalar_t *data() const { return d_.data(); }
    void resize(int row) {
        if (row != this->row()) d_.assign((size_t)(row > 0 ? row : 0), 0.0);
    }

private:
    std::vector<scalar_t> d_;
};

enum KernelId { K1x = IBIM_K1X, K1y, K1z, K2, K3xx, K3yy, K3zz, K3xy, K3yz, K3zx, K3yx, K3zy, K3xz, K4x, K4y, K4z };

class kernel {
public:
    std::function<scalar_t(point &, point &)> eval;  // not evaluable on the GPU; kept for the interface
    int kernelId = -1;                               // KernelId of the PB kernel this FMM applies
    ibim_params params{};                            // kappa, eps_in, eps_out, vacant radius

    kernel();
    ~kernel();
    kernel(const kernel &) = delete;
    kernel &operator=(const kernel &) = delete;

    // builds the octree and lists on the GPU (bbfmm.h:220-242); copies the
    // points and the charges like the reference
    void initialize(index_t nChebyshev, std::vector<point> &source, std::vector<point> &target, Vector &charge,
                    index_t nSource, index_t nTarget, index_t rank, index_t maxLevel);
    // potential(t) = sum_s K(source_s, target_t) charge(s) (bbfmm.h:244-270)
    void run(Vector &potential);

private:
    ibim_plan *plan_ = nullptr;
    bool same_targets_ = true;
    std::vector<double> tgt_;
    Vector charge_;
    index_t nTarget_ = 0;
};

}  // namespace bbfmm
