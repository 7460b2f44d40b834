// levelset.h -- the reference's geometry interface (levelset.h:12-72); the
// level set itself is computed and kept on the GPU (ibim_surface_*).
#pragma once

#include <vector>

#include "config.h"
#include "ibim.hpp"
#include "molecule.h"

// grid dimensions only: the values live in HBM during the pipeline
class Grid {
public:
    int Nx, Ny, Nz;
    Grid(int nx, int ny, int nz) : Nx(nx), Ny(ny), Nz(nz) {}
};

class levelset {
public:
    double dx, sx, sy, sz;
    int Nx, Ny, Nz;
    int bandwidth;
    int shift = 7;      // WENO5 stencil width
    int thickness;      // lvlset_thickness
    double thres;       // lvlset_thres
    levelset(int nx, int ny, int nz, int band, double sx, double sy, double sz, double dx, Config &cfg);
};

// tube quadrature nodes (levelset.h:62-72); the arrays are in the physical
// length unit (nodes / s, weights * s * (dx/s)^3, as electric.cpp:17-41)
class Surface {
public:
    // expand + evolve + flip + reinitialise x2 + Surface (main.cpp:47-62)
    Surface(levelset &ls, Molecule &mol, double probe, double reinit_time, double scale);
    // Surface of a level set given on the grid
    Surface(levelset &ls, const std::vector<double> &phi, double scale);
    ~Surface();
    Surface(const Surface &) = delete;
    Surface &operator=(const Surface &) = delete;

    ibim_surface *handle = nullptr;
    ibim_surface_info info{};
    std::vector<double> nodes, normals, weight;  // host copies, filled by fetch()
    void fetch();
    void output(const std::string &file);
};
