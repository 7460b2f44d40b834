/*
 * ibim_oracle.h -- CPU restatement of the reference IBIM Poisson-Boltzmann
 * FMM matvec path (GaZ3ll3/ibim-levelset), used ONLY as test infrastructure:
 * the parity checker for the HIP library and the timed CPU baseline.
 * Nothing in the product links or calls this code.
 *
 * What it restates (reference file:line):
 *   octree + interaction lists ...... bbfmm.cpp:8-280, bbfmm.h:24-198
 *   Chebyshev black-box FMM ......... bbfmm.h:200-623
 *   16 PB kernel closures ........... electric.cpp:67-291
 *   matvec dmapping ................. electric.cpp:294-388
 *   RHS load ........................ electric.cpp:395-413
 *   polarizeMap + energy ............ electric.cpp:432-505
 *   restarted GMRES ................. gmres.cpp:10-164
 *   level set + tube nodes .......... levelset.cpp:91-586, electric.cpp:17-41
 *                                     (ibim_levelset_oracle.c)
 *
 * Pinning (how each function is checked without running the reference,
 * which needs an external CBLAS library this image lacks):
 *   Chebyshev nodes/polys/S_p/R[8] -- closed forms and polynomial
 *     reproduction (tests/test_oracle_cheb.py);
 *   tree -- partition/containment/tie-breaking invariants and list
 *     symmetries (tests/test_oracle_tree.py);
 *   FMM engine -- with the polynomial surrogate kernels (mode 1) the
 *     Chebyshev FMM is exact, so it must equal the direct O(N^2) sum to
 *     rounding; this fails for any dropped/double-counted interaction,
 *     wrong transfer index or sign (tests/test_oracle_fmm.py);
 *   kernels -- finite differences of the Laplace/Yukawa Green's functions,
 *     kappa=0 and eps_in=eps_out special cases (tests/test_oracle_kernels.py);
 *   matvec -- FMM vs dense O(N^2) operator at p=6..8, A = I when
 *     eps_in = eps_out and kappa = 0;
 *   GMRES -- small dense systems against LAPACK (numpy.linalg.solve);
 *   whole solve -- the Born sphere energy;
 *   level set -- WENO5 derivatives exact on quadratics, Godunov norm of
 *     plane and sphere distance functions, evolve of a sphere moves it by
 *     the final time, reinitialisation keeps a distance function, Surface
 *     of a plane and a sphere against closed forms (tests/test_oracle_levelset.py).
 */
#ifndef IBIM_ORACLE_H
#define IBIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids, in the order of electric.cpp:299 */
enum {
    ORC_K1X = 0, ORC_K1Y, ORC_K1Z, ORC_K2,
    ORC_K3XX, ORC_K3YY, ORC_K3ZZ, ORC_K3XY, ORC_K3YZ, ORC_K3ZX,
    ORC_K3YX, ORC_K3ZY, ORC_K3XZ,
    ORC_K4X, ORC_K4Y, ORC_K4Z,
    ORC_NKERNEL
};

typedef struct {
    double kappa;          /* solvent_kappa */
    double eps_in;         /* solvent_dI */
    double eps_out;        /* solvent_dE */
    double vacant_radius;  /* tau * h  (electric.cpp:58) */
    int mode;              /* 0: PB kernels; 1: polynomial surrogate (test) */
} orc_kparams;

typedef struct {
    int p;                 /* fmm_np */
    int max_point;         /* fmm_maxpoint */
    int max_level;         /* fmm_maxlevel */
    orc_kparams k;
} orc_fmm_params;

typedef struct {
    int parent;
    int child[8];
    int level;
    int slot;              /* child slot in the parent (nodeIndex), root 0 */
    double c[3], r[3];     /* centre and per-axis radius */
    int nsrc, ntgt;
    int *src, *tgt;        /* original indices, ascending */
    int leaf, empty;
    int nu, nv, nw, nx;
    int *u, *v, *w, *x;    /* interaction lists (node ids) */
    int cu, cv, cw, cx;    /* capacities */
} orc_node;

typedef struct {
    int nnodes, cap;
    orc_node *nodes;
    int nsrc, ntgt;
    double *src;           /* 3*nsrc AoS copy */
    double *tgt;           /* 3*ntgt AoS copy */
    int max_point, max_level;
    int depth;             /* deepest leaf level */
} orc_tree;

/* tree (bbfmm.cpp:8-280) */
int  orc_tree_build(orc_tree *t, const double *src, int nsrc, const double *tgt, int ntgt,
                    int max_point, int max_level);
void orc_tree_free(orc_tree *t);

/* flat export for tests: arrays sized nnodes (centre/radius 3*nnodes) */
int  orc_tree_nnodes(const orc_tree *t);
void orc_tree_export(const orc_tree *t, int *level, int *parent, int *slot, int *leaf, int *empty,
                     int *nsrc, int *ntgt, double *centre, double *radius,
                     int *nu, int *nv, int *nw, int *nx);
/* list of kind 0=U 1=V 2=W 3=X of node id into out (capacity cap); returns its length */
int  orc_tree_list(const orc_tree *t, int id, int kind, int *out, int cap);
/* sources/targets of node id (original indices); returns count */
int  orc_tree_points(const orc_tree *t, int id, int which, int *out, int cap);

/* Chebyshev building blocks (bbfmm.h:273-433) */
void orc_cheb_nodes(int p, double *x);
void orc_cheb_poly(int p, int n, const double *x, double *T /* n x p col-major */);
void orc_sp(int p, int n, const double *x, double *S /* n x p col-major */);
void orc_transfer_R(int p, double *R /* 8 blocks of r x r col-major */);

/* one kernel closure: a = source, b = target (electric.cpp:67-291) */
double orc_kernel_eval(int kid, const orc_kparams *kp, const double *a, const double *b);

/* one black-box FMM (bbfmm::kernel::initialize + run, bbfmm.h:220-270) on a
 * built tree; charge in source order, potential in target order. */
int orc_fmm_run(const orc_tree *t, int p, int kid, const orc_kparams *kp,
                const double *charge, double *potential);

/* direct O(nsrc*ntgt) sum of the same kernel, for brute-force checks */
void orc_direct(int kid, const orc_kparams *kp, const double *src, int nsrc,
                const double *tgt, int ntgt, const double *charge, double *potential);

/* the matvec y = A x, x = [phi; psi] (electric.cpp:294-388).
 * xyz, nrm: 3n AoS; w: n.  share_tree=0 builds one tree per kernel like the
 * reference; 1 builds one tree and reuses it (same result, faster). */
int orc_dmapping(int n, const double *xyz, const double *nrm, const double *w,
                 const orc_fmm_params *fp, const double *x, double *y, int share_tree);

/* dense O(N^2) version of the same operator (every kernel summed directly) */
void orc_dmapping_direct(int n, const double *xyz, const double *nrm, const double *w,
                         const orc_kparams *kp, const double *x, double *y);

/* RHS `load` (electric.cpp:395-413) */
void orc_rhs(int n, const double *xyz, const double *nrm, int m, const double *atom_xyz,
             const double *atom_q, double eps_in, double eps_out, double *b);

/* polarizeMap (electric.cpp:442-485): reaction potential at the m atoms */
int orc_polarize(int n, const double *xyz, const double *nrm, const double *w, int m,
                 const double *atom_xyz, const orc_fmm_params *fp, const double *x,
                 double *phi_reac);

/* restarted GMRES(m) with MGS and Givens (gmres.cpp:50-164).
 * Returns 0 on convergence, 1 otherwise; *iters = inner iterations used,
 * *resid = final relative residual, *nmatvec = operator applications. */
typedef void (*orc_matvec_fn)(void *ctx, const double *x, double *y);
int orc_gmres(orc_matvec_fn A, void *ctx, int n, const double *b, double *x, int m, int maxit,
              double tol, int *iters, double *resid, int *nmatvec);

/* GMRES on a dense col-major n x n matrix (for tests) */
int orc_gmres_dense(int n, const double *Amat, const double *b, double *x, int m, int maxit,
                    double tol, int *iters, double *resid);

/* GMRES with the FMM matvec as operator; share_tree as in orc_dmapping */
int orc_solve(int n, const double *xyz, const double *nrm, const double *w,
              const orc_fmm_params *fp, const double *b, double *x, int m, int maxit,
              double tol, int share_tree, int *iters, double *resid, int *nmatvec);

/* ---- geometry pre-step (ibim_levelset_oracle.c) ----------------------------
 * The level set of the solvent-excluded surface and the tube nodes:
 * expand, evolve, reinitialize, Surface (levelset.cpp:91-586, main.cpp:43-62).
 * Grid index I = i*ny*nz + j*nz + k, coordinates lo + (i,j,k)*h. */
typedef struct {
    int dims[3];
    double lo[3];
    double h;
    double probe;        /* probe radius; also the evolve's final time at speed 1 */
    int band;            /* expand window margin in cells */
    int thickness;       /* tube half width in cells */
    double thres;        /* reinitialisation gradient tolerance */
    double reinit_time;  /* final time of each reinitialisation */
    double scale;        /* nodes / scale, weights * scale * (h / scale)^3 */
} orc_lsparams;

typedef struct {
    int evolve_steps, reinit_steps[2];
    double quality[2];   /* mean ||grad phi| - 1| and bad fraction in the tube, after the last pass */
    int64_t inclusions, n_nodes;
} orc_lsinfo;

/* the pieces, for pinning: WENO5 one-sided derivatives at (i, j, k)
 * (setGradient), the Godunov norm getNorm(A, B), and one evolve (mode 0) or
 * reinitialize (mode 1) of g in place, returning the steps taken */
void orc_ls_gradient(const orc_lsparams *p, const double *g, int i, int j, int k, double Dup[3], double Dun[3]);
double orc_ls_norm(const double A[3], const double B[3]);
int orc_ls_run(const orc_lsparams *p, int mode, double *g, const double *phi0, double final_t, double vel, double cfl,
               double quality[2]);
/* Surface of phi (not modified); xyz/nrm (3 per node) and w are filled when
 * xyz is non-null and hold cap nodes.  Returns the node count, -1 if cap is short. */
int64_t orc_surface_from_grid(const orc_lsparams *p, const double *phi, double *xyz, double *nrm, double *w,
                              int64_t cap, int64_t *inclusions);
/* the whole pre-step from atoms on the grid p->dims/p->lo, the grid starting
 * at `init`; g_out (optional) receives the level set before Surface */
int64_t orc_surface_from_atoms(const orc_lsparams *p, int m, const double *axyz, const double *radii,
                               double init, double *xyz, double *nrm, double *w, int64_t cap, orc_lsinfo *info,
                               double *g_out);

#ifdef __cplusplus
}
#endif
#endif
