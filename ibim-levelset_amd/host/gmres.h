// gmres.h -- the reference's solver interface (gmres.h:20-21).
#pragma once

#include <functional>

#include "bbfmm.h"

// Restarted GMRES(m) on a host operator A: modified Gram-Schmidt Arnoldi,
// Givens rotations, residual |b - Ax| / |b| (gmres.cpp:50-164).  Returns 0 on
// convergence, 1 otherwise; x holds the initial guess and the solution.
int GMRES(const std::function<bbfmm::Vector(bbfmm::Vector &)> A, bbfmm::Vector &x, bbfmm::Vector &b, int m,
          int max_iter, double tol);

// The same solve with the operator of a GPU plan: the Krylov basis, the
// Hessenberg matrix and the rotations stay in HBM (ibim_gmres).
int GMRES(ibim_plan *plan, bbfmm::Vector &x, bbfmm::Vector &b, int m, int max_iter, double tol,
          ibim_solve_info *info = nullptr);
