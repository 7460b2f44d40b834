// molecule.h -- atoms of a PQR file (molecule.h:10-126 of the reference).
#pragma once

#include <string>
#include <vector>

class Molecule {
public:
    std::vector<double> centers;  // 3 per atom
    std::vector<double> radii;
    std::vector<double> charges;
    double center[3] = {0, 0, 0};
    double radius = 0;
    int N = 0;

    // ATOM lines: x y z = tokens 5-7, charge = token 8, radius = token 9
    bool load(const std::string &file);
    // bounding box of the atom spheres: centre and largest half extent
    void getCenter();
    // move the centre to the origin and scale so the half extent is `range`;
    // returns the scale s (molecule.h:112-125)
    double centralize(double range);
    // seeded synthetic globule (ibim_synthetic_globule)
    void synthetic(int atoms, unsigned long long seed, double confine);
};
