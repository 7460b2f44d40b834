/* Prints the size and field offsets of every struct include/ibim.h declares,
 * one "struct size" line and one "struct.field offset" line each, for the
 * ctypes layout check in tests/test_abi.py.  Host C only. */
#include <stddef.h>
#include <stdio.h>

#include "ibim.h"

#define S(t) printf("%s %zu\n", #t, sizeof(t))
#define F(t, f) printf("%s.%s %zu\n", #t, #f, offsetof(t, f))

int main(void) {
    S(ibim_params);
    F(ibim_params, p); F(ibim_params, max_points); F(ibim_params, max_level); F(ibim_params, kappa);
    F(ibim_params, eps_in); F(ibim_params, eps_out); F(ibim_params, vacant_radius); F(ibim_params, precision);
    F(ibim_params, kernel_mode); F(ibim_params, inner_order);
    S(ibim_tree_stats);
    F(ibim_tree_stats, n_points); F(ibim_tree_stats, n_nodes); F(ibim_tree_stats, n_leaves);
    F(ibim_tree_stats, n_empty_leaves); F(ibim_tree_stats, depth); F(ibim_tree_stats, n_u);
    F(ibim_tree_stats, n_v); F(ibim_tree_stats, n_w); F(ibim_tree_stats, n_x); F(ibim_tree_stats, p2p_pairs);
    F(ibim_tree_stats, m2l_pairs); F(ibim_tree_stats, v_pairs); F(ibim_tree_stats, x_pairs);
    F(ibim_tree_stats, p2p_pairs_u);
    S(ibim_solve_info);
    F(ibim_solve_info, iterations); F(ibim_solve_info, matvecs); F(ibim_solve_info, rel_residual);
    F(ibim_solve_info, converged); F(ibim_solve_info, outer_steps);
    S(ibim_stage_report);
    F(ibim_stage_report, ms); F(ibim_stage_report, launches); F(ibim_stage_report, valu_flops);
    F(ibim_stage_report, mfma_flops); F(ibim_stage_report, bytes);
    S(ibim_geom_params);
    F(ibim_geom_params, h); F(ibim_geom_params, lo); F(ibim_geom_params, dims); F(ibim_geom_params, pad);
    F(ibim_geom_params, probe); F(ibim_geom_params, band); F(ibim_geom_params, thickness);
    F(ibim_geom_params, thres); F(ibim_geom_params, reinit_time); F(ibim_geom_params, scale);
    F(ibim_geom_params, init);
    S(ibim_surface_info);
    F(ibim_surface_info, n_nodes); F(ibim_surface_info, dims); F(ibim_surface_info, lo); F(ibim_surface_info, h);
    F(ibim_surface_info, inclusions); F(ibim_surface_info, evolve_steps); F(ibim_surface_info, reinit_steps);
    F(ibim_surface_info, quality); F(ibim_surface_info, area);
    S(ibim_request);
    F(ibim_request, n_atoms); F(ibim_request, atom_xyz); F(ibim_request, radii); F(ibim_request, charges);
    F(ibim_request, geom); F(ibim_request, fmm); F(ibim_request, restart); F(ibim_request, maxit);
    F(ibim_request, tol);
    S(ibim_result);
    F(ibim_result, status); F(ibim_result, energy); F(ibim_result, n_points); F(ibim_result, info);
    F(ibim_result, t_submit); F(ibim_result, t_start); F(ibim_result, t_end); F(ibim_result, footprint_est);
    F(ibim_result, footprint);
    return 0;
}
