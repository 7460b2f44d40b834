// levelset <input.cfg> -- the reference's executable (main.cpp:14-79):
// molecule -> solvent-excluded level set -> tube nodes -> PB solve -> energy.
//
// Reference mode (no grid_h key): the molecule is scaled so its half extent is
// 240 units inside a [-300,300]^3 grid of grid_size cells, as main.cpp:33-45.
// Angstrom mode (grid_h = spacing in A): the grid is fitted around the atoms
// with margin grid_pad (default 12 h) and no scaling.  Either way the
// molecule is a PQR file (pqr_file) or, with synthetic_atoms = M, the seeded
// synthetic globule (synthetic_seed, synthetic_radius).
#include <chrono>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>

#include "config.h"
#include "electric.h"
#include "levelset.h"
#include "molecule.h"

int main(int argc, char *argv[]) {
    if (argc <= 1) {
        std::cout << "USE " << argv[0] << " [--dry-run] PATH_OF_CONFIG_FILE " << std::endl;
        return 0;
    }
    // --dry-run: read the options and the molecule, print the grid, stop
    // before any GPU work
    const bool dry = std::string(argv[1]) == "--dry-run";
    const char *cfg_path = argv[dry && argc > 2 ? 2 : 1];
    try {
        Config cfg;
        std::ifstream in(cfg_path);
        if (!in.is_open()) {
            std::cerr << "cannot open " << cfg_path << std::endl;
            return 2;
        }
        cfg.parse(in);
        cfg.print();

        Molecule mol;
        if (cfg.has("synthetic_atoms")) {
            mol.synthetic(cfg.geti("synthetic_atoms", 500), (unsigned long long)cfg.getd("synthetic_seed", 500),
                          cfg.getd("synthetic_radius", 20.0));
        } else if (!mol.load(cfg.gets("pqr_file", ""))) {
            std::cerr << "cannot read atoms from pqr_file " << cfg.gets("pqr_file", "") << std::endl;
            return 2;
        }
        const double probe_A = 1.4;
        double s = 1.0, dx, lo[3];
        int dims[3];
        double reinit_time;
        if (cfg.has("grid_h")) {
            dx = cfg.getd("grid_h", 0.25);
            const double pad = cfg.getd("grid_pad", 12 * dx);
            double bl[3] = {1e300, 1e300, 1e300}, bh[3] = {-1e300, -1e300, -1e300};
            for (int i = 0; i < mol.N; ++i)
                for (int d = 0; d < 3; ++d) {
                    bl[d] = std::min(bl[d], mol.centers[3 * i + d] - mol.radii[i] - probe_A);
                    bh[d] = std::max(bh[d], mol.centers[3 * i + d] + mol.radii[i] + probe_A);
                }
            for (int d = 0; d < 3; ++d) {
                lo[d] = bl[d] - pad;
                dims[d] = (int)std::ceil((bh[d] + pad - lo[d]) / dx) + 1;
            }
            reinit_time = cfg.getd("reinit_time", 26.0 * dx);
            std::cout << std::setw(15) << "ATOMS" << " " << std::setw(8) << mol.N << std::endl;
        } else {
            s = mol.centralize(240.0);
            mol.getCenter();
            const int size = cfg.geti("grid_size", 256);
            dx = 600.0 / size;
            lo[0] = lo[1] = lo[2] = -300.0;
            dims[0] = dims[1] = dims[2] = size;
            reinit_time = cfg.getd("reinit_step", 60);
        }
        std::cout << std::setw(15) << "h" << " " << std::setw(8) << dx / s << " Angstroms" << std::endl;
        if (dry) {
            double q = 0;
            for (double c : mol.charges) q += c;
            std::cout << std::setw(15) << "GRID" << " " << dims[0] << " x " << dims[1] << " x " << dims[2] << " from "
                      << lo[0] << " " << lo[1] << " " << lo[2] << std::endl;
            std::cout << std::setw(15) << "NET CHARGE" << " " << q << std::endl;
            return 0;
        }
        levelset ls(dims[0], dims[1], dims[2], 8, lo[0], lo[1], lo[2], dx, cfg);
        Grid g(dims[0], dims[1], dims[2]);
        auto t0 = std::chrono::steady_clock::now();
        Surface surf(ls, mol, probe_A * s, reinit_time, s);
        auto t1 = std::chrono::steady_clock::now();
        std::cout << std::setw(15) << "SURFACE" << " " << std::setprecision(5) << std::setw(8)
                  << std::chrono::duration<double>(t1 - t0).count() << " seconds" << std::endl;
        electric(g, ls, surf, mol, s, cfg);
        auto t2 = std::chrono::steady_clock::now();
        std::cout << std::setw(15) << "SOLVE" << " " << std::setprecision(5) << std::setw(8)
                  << std::chrono::duration<double>(t2 - t1).count() << " seconds" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "levelset: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
