#include "bbfmm.h"

#include <cstring>

namespace bbfmm {

kernel::kernel() {
    params.p = 4;
    params.max_points = 64;
    params.max_level = 10;
    params.eps_in = 1.0;
    params.eps_out = 80.0;
    params.precision = IBIM_PREC_FP64;
    params.kernel_mode = IBIM_KERNEL_PB;
}

kernel::~kernel() { ibim_plan_destroy(plan_); }

void kernel::initialize(index_t nChebyshev, std::vector<point> &source, std::vector<point> &target,
                        Vector &charge, index_t nSource, index_t nTarget, index_t rank, index_t maxLevel) {
    if (kernelId < 0 || kernelId > 15)
        throw ibim::Error("bbfmm::kernel: set kernelId (a black-box closure cannot run on the GPU)");
    params.p = nChebyshev;
    params.max_points = rank;
    params.max_level = maxLevel;
    std::vector<double> xyz(3 * (size_t)nSource), nrm(3 * (size_t)nSource, 0.0), w((size_t)nSource, 1.0);
    for (index_t i = 0; i < nSource; ++i) {
        xyz[3 * i] = source[i].x;
        xyz[3 * i + 1] = source[i].y;
        xyz[3 * i + 2] = source[i].z;
    }
    same_targets_ = (nTarget == nSource);
    for (index_t i = 0; same_targets_ && i < nTarget; ++i)
        same_targets_ = target[i].x == source[i].x && target[i].y == source[i].y && target[i].z == source[i].z;
    tgt_.assign(3 * (size_t)nTarget, 0.0);
    for (index_t i = 0; i < nTarget; ++i) {
        tgt_[3 * i] = target[i].x;
        tgt_[3 * i + 1] = target[i].y;
        tgt_[3 * i + 2] = target[i].z;
    }
    nTarget_ = nTarget;
    charge_ = charge;
    ibim_plan_destroy(plan_);
    plan_ = nullptr;
    ibim::check(ibim_plan_create(&params, nSource, xyz.data(), nrm.data(), w.data(), &plan_), "ibim_plan_create");
}

void kernel::run(Vector &potential) {
    if (!plan_) throw ibim::Error("bbfmm::kernel::run before initialize");
    potential = Vector(nTarget_);
    if (same_targets_)
        ibim::check(ibim_kernel_fmm(plan_, kernelId, charge_.data(), potential.data()), "ibim_kernel_fmm");
    else
        ibim::check(ibim_kernel_fmm_targets(plan_, kernelId, nTarget_, tgt_.data(), charge_.data(), potential.data()),
                    "ibim_kernel_fmm_targets");
}

}  // namespace bbfmm
