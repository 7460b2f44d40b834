#include "levelset.h"

#include <algorithm>
#include <fstream>
#include <iomanip>
#include <iostream>

levelset::levelset(int nx, int ny, int nz, int band, double _sx, double _sy, double _sz, double _dx, Config &cfg)
    : dx(_dx), sx(_sx), sy(_sy), sz(_sz), Nx(nx), Ny(ny), Nz(nz), bandwidth(band) {
    thickness = cfg.geti("lvlset_thickness", 2);
    thres = cfg.getd("lvlset_thres", 1e-4);
}

static ibim_geom_params geom_of(const levelset &ls, double scale) {
    ibim_geom_params gp{};
    gp.h = ls.dx;
    gp.lo[0] = ls.sx;
    gp.lo[1] = ls.sy;
    gp.lo[2] = ls.sz;
    gp.dims[0] = ls.Nx;
    gp.dims[1] = ls.Ny;
    gp.dims[2] = ls.Nz;
    gp.band = ls.bandwidth;
    gp.thickness = ls.thickness;
    gp.thres = ls.thres;
    gp.scale = scale;
    // the grid starts at -2 * size before expand (main.cpp:43)
    gp.init = -2.0 * std::max(ls.Nx, std::max(ls.Ny, ls.Nz));
    return gp;
}

static void report(const ibim_surface_info &inf) {
    std::cout << std::setw(15) << "QUALITY" << " " << std::setprecision(2) << std::setw(8) << inf.quality[0]
              << " & " << inf.quality[1] << std::endl;
    std::cout << std::setw(15) << "INCLUSION" << " " << std::setw(8) << inf.inclusions << std::endl;
}

Surface::Surface(levelset &ls, Molecule &mol, double probe, double reinit_time, double scale) {
    ibim_geom_params gp = geom_of(ls, scale);
    gp.probe = probe;
    gp.reinit_time = reinit_time;
    ibim::check(ibim_surface_from_atoms(&gp, mol.N, mol.centers.data(), mol.radii.data(), &handle),
                "ibim_surface_from_atoms");
    ibim::check(ibim_surface_get_info(handle, &info), "ibim_surface_get_info");
    report(info);
}

Surface::Surface(levelset &ls, const std::vector<double> &phi, double scale) {
    ibim_geom_params gp = geom_of(ls, scale);
    ibim::check(ibim_surface_from_grid(&gp, phi.data(), &handle), "ibim_surface_from_grid");
    ibim::check(ibim_surface_get_info(handle, &info), "ibim_surface_get_info");
}

Surface::~Surface() { ibim_surface_destroy(handle); }

void Surface::fetch() {
    const size_t n = (size_t)info.n_nodes;
    nodes.resize(3 * n);
    normals.resize(3 * n);
    weight.resize(n);
    ibim::check(ibim_surface_copy(handle, nodes.data(), normals.data(), weight.data()), "ibim_surface_copy");
}

void Surface::output(const std::string &file) {
    if (nodes.empty()) fetch();
    std::ofstream out(file);
    for (size_t i = 0; i < weight.size(); ++i)
        out << nodes[3 * i] << " " << nodes[3 * i + 1] << " " << nodes[3 * i + 2] << "\n";
}
