// electric.h -- the PB solve (electric.h:13 of the reference).
#pragma once

#include "config.h"
#include "levelset.h"
#include "molecule.h"

// Builds the GPU plan over the surface, forms the right-hand side from the
// atoms (centres / rescale), solves with GMRES on the device, writes
// potent_file (2N values) and energy_file (reaction potential per atom), and
// prints the polarized energy (1/2) sum q phi_reac (electric.cpp:9-507).
void electric(Grid &g, levelset &ls, Surface &surf, Molecule &mol, double rescale, Config &cfg);
