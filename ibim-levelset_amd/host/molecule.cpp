#include "molecule.h"

#include <algorithm>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "ibim.hpp"

bool Molecule::load(const std::string &file) {
    std::ifstream in(file);
    if (!in.is_open()) return false;
    for (std::string line; std::getline(in, line);) {
        if (line.find("ATOM") != 0) continue;
        std::stringstream ss(line);
        std::vector<std::string> tok;
        for (std::string b; ss >> b;) tok.push_back(b);
        if (tok.size() < 10) continue;
        centers.push_back(std::stod(tok[5]));
        centers.push_back(std::stod(tok[6]));
        centers.push_back(std::stod(tok[7]));
        charges.push_back(std::stod(tok[8]));
        radii.push_back(std::stod(tok[9]));
    }
    N = (int)radii.size();
    if (N > 0) getCenter();
    return N > 0;
}

void Molecule::getCenter() {
    double lo[3], hi[3];
    for (int j = 0; j < 3; ++j) {
        lo[j] = centers[j] - radii[0];
        hi[j] = centers[j] + radii[0];
    }
    for (int i = 1; i < N; ++i)
        for (int j = 0; j < 3; ++j) {
            hi[j] = std::max(hi[j], centers[3 * i + j] + radii[i]);
            lo[j] = std::min(lo[j], centers[3 * i + j] - radii[i]);
        }
    for (int j = 0; j < 3; ++j) center[j] = (lo[j] + hi[j]) * 0.5;
    radius = std::max(0.5 * (hi[0] - lo[0]), std::max(0.5 * (hi[1] - lo[1]), 0.5 * (hi[2] - lo[2])));
}

double Molecule::centralize(double range) {
    const double s = range / radius;
    for (int i = 0; i < N; ++i) {
        for (int j = 0; j < 3; ++j) centers[3 * i + j] = s * (centers[3 * i + j] - center[j]);
        radii[i] = s * radii[i];
    }
    std::cout << std::setw(15) << "ATOMS" << " " << std::setw(8) << N << std::endl;
    return s;
}

void Molecule::synthetic(int atoms, unsigned long long seed, double confine) {
    N = atoms;
    centers.assign(3 * (size_t)atoms, 0.0);
    radii.assign(atoms, 0.0);
    charges.assign(atoms, 0.0);
    ibim::check(ibim_synthetic_globule(atoms, seed, confine, centers.data(), radii.data(), charges.data()),
                "ibim_synthetic_globule");
    getCenter();
}
