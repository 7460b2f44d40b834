#include "electric.h"

#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

#include "bbfmm.h"
#include "gmres.h"

void electric(Grid &, levelset &ls, Surface &surf, Molecule &mol, double rescale, Config &cfg) {
    const int64_t N = surf.info.n_nodes;
    std::cout << std::setw(15) << "POINTS NUM" << " " << std::setw(8) << N << std::endl;
    std::cout << std::setw(15) << "AREA APPROX" << " " << std::setw(8) << surf.info.area << " A^2" << std::fixed
              << std::endl;
    const double dx = ls.dx / rescale;
    ibim_params prm{};
    prm.p = cfg.geti("fmm_np", 4);
    prm.max_points = cfg.geti("fmm_maxpoint", 64);
    prm.max_level = cfg.geti("fmm_maxlevel", 10);
    prm.kappa = cfg.getd("solvent_kappa", 0.0);
    prm.eps_out = cfg.getd("solvent_dE", 80.0);
    prm.eps_in = cfg.getd("solvent_dI", 1.0);
    prm.vacant_radius = cfg.getd("tau", 1.0) * dx;  // electric.cpp:58
    const std::string prec = cfg.gets("precision", "fp64");
    prm.precision = prec == "mixed" ? IBIM_PREC_MIXED : prec == "mixed16" ? IBIM_PREC_MIXED_F16 : IBIM_PREC_FP64;
    prm.kernel_mode = IBIM_KERNEL_PB;
    if (prm.precision != IBIM_PREC_FP64) prm.inner_order = cfg.geti("inner_order", 0);

    const double *dxyz, *dnrm, *dw;
    ibim::check(ibim_surface_device(surf.handle, &dxyz, &dnrm, &dw), "ibim_surface_device");
    ibim_plan *plan = nullptr;
    ibim::check(ibim_plan_create_device(&prm, N, dxyz, dnrm, dw, &plan), "ibim_plan_create_device");

    std::vector<double> atoms(3 * (size_t)mol.N);
    for (size_t i = 0; i < atoms.size(); ++i) atoms[i] = mol.centers[i] / rescale;
    bbfmm::Vector load(2 * (int)N), start(2 * (int)N);
    ibim::check(ibim_rhs(plan, mol.N, atoms.data(), mol.charges.data(), load.data()), "ibim_rhs");
    ibim_solve_info info{};
    GMRES(plan, start, load, cfg.geti("gmres_restart", 40), cfg.geti("gmres_maxiter", 100),
          cfg.getd("gmres_tol", 1e-5), &info);

    {
        std::ofstream out(cfg.gets("potent_file", "out.potent"));
        for (int i = 0; i < 2 * N; ++i) out << start(i) << "\n";
    }
    std::vector<double> phi((size_t)mol.N);
    ibim::check(ibim_polarize(plan, mol.N, atoms.data(), start.data(), phi.data()), "ibim_polarize");
    {
        std::ofstream out(cfg.gets("energy_file", "out.energy"));
        for (double v : phi) out << v << "\n";
    }
    double energy = 0.0;
    for (int i = 0; i < mol.N; ++i) energy += phi[i] * mol.charges[i];
    energy *= 0.5;
    std::cout << "polarized energy: " << std::setw(20) << std::scientific << energy << std::fixed << std::endl;
    ibim_plan_destroy(plan);
}
