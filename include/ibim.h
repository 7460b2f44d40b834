/*
 * ibim.h -- C ABI of the MI355X-native IBIM Poisson-Boltzmann FMM matvec.
 *
 * One plan = one surface point set (the level-set tube nodes with normals
 * and weights).  Creating a plan builds the reference octree and its U/V/W/X
 * interaction lists on the GPU once; every matvec, FMM, RHS, energy and GMRES
 * call then reuses that device-resident state.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Functions without a `_device` suffix
 *     take host arrays and return when the results are in them.  `_device`
 *     variants take device pointers (HBM) and are ordered on the plan's
 *     stream; call ibim_plan_sync() before reading their results.  Device
 *     inputs must be complete when a call is made (synchronise the stream
 *     that wrote them first): the plan's stream does not wait for others.
 *   - Points are AoS: xyz[3*i+0..2]; normals the same; one weight per point.
 *   - Return 0 (IBIM_OK) on success, a negative IBIM_ERR_* otherwise;
 *     ibim_last_error() describes the last failure of the calling thread.
 *   - Plans are independent: different plans may be used from different
 *     threads.  One plan must not be used by two threads at once.
 *
 * Reference interfaces each entry point replaces (GaZ3ll3/ibim-levelset):
 *   ibim_plan_create      bbfmm::kernel::initialize, bbfmm.h:220-242, with the
 *                         tree build tree::populate, bbfmm.cpp:8-41, done once
 *                         per geometry instead of once per kernel per matvec
 *   ibim_kernel_fmm       bbfmm::kernel::run, bbfmm.h:244-270, for one of the
 *                         16 kernel closures of electric.cpp:67-291
 *   ibim_matvec           the dmapping lambda, electric.cpp:294-388, called by
 *                         GMRES through FullMap, electric.cpp:390-392
 *   ibim_rhs              the `load` loop, electric.cpp:395-413
 *   ibim_polarize         polarizeMap, electric.cpp:442-485
 *   ibim_gmres            GMRES(A, x, b, m, max_iter, tol), gmres.h:20-21,
 *                         with A = the matvec of this plan
 *   ibim_tree_*           tree::output, bbfmm.cpp:283-302 (per-node dump)
 */
#ifndef IBIM_H
#define IBIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IBIM_OK 0
#define IBIM_ERR_ARG (-1)     /* invalid argument */
#define IBIM_ERR_HIP (-2)     /* HIP runtime failure (no device, launch error) */
#define IBIM_ERR_NOMEM (-3)   /* device allocation failed */
#define IBIM_ERR_STATE (-4)   /* call not valid for this plan */
#define IBIM_NOT_CONVERGED 1  /* ibim_gmres: iteration limit reached */

#define IBIM_PREC_FP64 0      /* everything in FP64 (the reference's arithmetic) */
#define IBIM_PREC_MIXED 1     /* FP32 far/near field inside FP64 refinement */
#define IBIM_PREC_MIXED_F16 2 /* as MIXED, V- and X-list M2L on FP16 MFMA with power-of-two
                                 scales per box and per operator row, FP32 accumulation
                                 (p <= 6; FP32 M2L above) */

#define IBIM_KERNEL_PB 0      /* the 16 PB kernels of electric.cpp:67-291 */
#define IBIM_KERNEL_POLY 1    /* polynomial surrogates (FMM-exact; for testing) */

/* kernel ids of ibim_kernel_fmm, in the order of electric.cpp:299 */
enum {
    IBIM_K1X = 0, IBIM_K1Y, IBIM_K1Z, IBIM_K2,
    IBIM_K3XX, IBIM_K3YY, IBIM_K3ZZ, IBIM_K3XY, IBIM_K3YZ, IBIM_K3ZX,
    IBIM_K3YX, IBIM_K3ZY, IBIM_K3XZ,
    IBIM_K4X, IBIM_K4Y, IBIM_K4Z
};

typedef struct ibim_plan ibim_plan;

typedef struct ibim_params {
    int p;                /* Chebyshev order, fmm_np (2..8) */
    int max_points;       /* leaf capacity, fmm_maxpoint (>= 1) */
    int max_level;        /* deepest tree level, fmm_maxlevel (0..20) */
    double kappa;         /* solvent_kappa, 1/Angstrom (>= 0) */
    double eps_in;        /* solvent_dI (> 0) */
    double eps_out;       /* solvent_dE (> 0) */
    double vacant_radius; /* tau * h, Angstrom: kernels regularised below it */
    int precision;        /* IBIM_PREC_* */
    int kernel_mode;      /* IBIM_KERNEL_* */
    int inner_order;      /* mixed forms: Chebyshev order of the operator the inner solves use
                             (2..p-1; 0: p).  The FP64 residual of the refinement keeps order
                             p, so the converged answer is still that of the order-p FP64
                             system.  No counterpart in the reference (FP64 GMRES only). */
} ibim_params;

typedef struct ibim_tree_stats {
    int64_t n_points;
    int64_t n_nodes;
    int64_t n_leaves;
    int64_t n_empty_leaves;
    int depth;            /* deepest leaf level */
    int64_t n_u, n_v, n_w, n_x;      /* total list lengths (all nodes) */
    int64_t p2p_pairs;    /* sum over target leaves of nT * sum_{U,W} nS */
    int64_t m2l_pairs;    /* non-empty V + X interactions into nodes with targets */
    int64_t v_pairs;      /* the V part of m2l_pairs */
    int64_t x_pairs;      /* the X part of m2l_pairs */
    int64_t p2p_pairs_u;  /* the U-list part of p2p_pairs (the rest is the W list) */
} ibim_tree_stats;

typedef struct ibim_solve_info {
    int iterations;       /* Arnoldi steps (inner iterations) */
    int matvecs;          /* operator applications incl. residual checks */
    double rel_residual;  /* final |b - Ax| / |b| estimate */
    int converged;        /* 1 if rel_residual < tol */
    int outer_steps;      /* FP64 refinement steps (mixed precision only) */
} ibim_solve_info;

/* ---- plan --------------------------------------------------------------- */
int  ibim_plan_create(const ibim_params *prm, int64_t n, const double *xyz, const double *normals,
                      const double *weights, ibim_plan **out);
int  ibim_plan_create_device(const ibim_params *prm, int64_t n, const double *d_xyz,
                             const double *d_normals, const double *d_weights, ibim_plan **out);
void ibim_plan_destroy(ibim_plan *plan);
int  ibim_plan_sync(ibim_plan *plan);
/* order the plan's work on a caller's HIP stream (hipStream_t passed as
 * void*; NULL: back to the plan's own stream).  Every later call is then
 * enqueued on that stream after whatever the caller queued there before it,
 * so device inputs need no synchronisation and later work on the stream
 * sees the results; the plan's earlier work is ordered before the switch.
 * The stream must outlive its use by the plan.  No counterpart in the
 * reference, which runs on the host. */
int  ibim_plan_set_stream(ibim_plan *plan, void *stream);
int  ibim_plan_set_precision(ibim_plan *plan, int precision);
/* device memory held by the plan, bytes */
int64_t ibim_plan_footprint(const ibim_plan *plan);

/* stage timing: HIP events recorded on the plan's stream around each stage
 * while enabled (the reference's RUN timing macro, utils.h:45-56, but on the
 * device); ibim_plan_stage_times waits for them and returns milliseconds and
 * counts per stage (IBIM_STAGE_COUNT entries), optionally resetting them. */
#define IBIM_STAGE_P2M 0
#define IBIM_STAGE_M2M 1
#define IBIM_STAGE_M2L 2      /* V list (GEMM) or all of M2L (direct form) */
#define IBIM_STAGE_L2L 3
#define IBIM_STAGE_P2P 4      /* near field + L2P */
#define IBIM_STAGE_MATVEC 5   /* one whole matvec */
#define IBIM_STAGE_ORTH 6     /* GMRES orthogonalisation */
#define IBIM_STAGE_RHS 7
#define IBIM_STAGE_POLARIZE 8
#define IBIM_STAGE_M2L_X 9    /* X list, direct, when the V list runs as GEMM */
#define IBIM_STAGE_MATVEC_MIXED 10 /* one mixed-precision matvec */
#define IBIM_STAGE_M2L_MIXED 11    /* V list of a mixed-precision matvec (counted here, not in M2L) */
#define IBIM_STAGE_P2P_MIXED 12    /* near field of a mixed-precision matvec (not in P2P): with FP16 M2L its
                                      two launches on the second stream (costliest units beside the upward
                                      pass, the rest beside the M2L reduction), else near field + L2P */
#define IBIM_STAGE_COUNT 13
int ibim_plan_profile(ibim_plan *plan, int enable);
/* record only the stages whose bit (1u << IBIM_STAGE_x) is set; 0 = off.
   Each recorded stage costs two events on the plan's stream per launch.
   No counterpart in the reference (which has no stage timers). */
int ibim_plan_profile_stages(ibim_plan *plan, unsigned stage_mask);
int ibim_plan_stage_times(ibim_plan *plan, double *ms, int64_t *launches, int reset);
/* the same, with the algorithmic work done inside the recorded intervals of
 * each stage, counted as the roofline model of the matvec counts it: 60
 * flops per near-field pair (vector ALU), 32 p^6 flops per V or X
 * interaction (matrix cores), and 8 (14 n + 8 p^3 B) bytes of points and
 * expansions per matvec (B non-empty boxes).  Only the matvec stages (whole
 * matvec, M2L, near field) carry work.  No counterpart in the reference. */
#define IBIM_STAGE_SLOTS 16
typedef struct ibim_stage_report {
    double ms[IBIM_STAGE_SLOTS];          /* summed interval lengths */
    int64_t launches[IBIM_STAGE_SLOTS];   /* recorded intervals */
    double valu_flops[IBIM_STAGE_SLOTS];
    double mfma_flops[IBIM_STAGE_SLOTS];
    double bytes[IBIM_STAGE_SLOTS];
} ibim_stage_report;
int ibim_plan_stage_report(ibim_plan *plan, ibim_stage_report *out, int reset);

/* ---- the hot path -------------------------------------------------------- */
/* y = A x, x = [phi (n); psi (n)], exactly the reference dmapping */
int ibim_matvec(ibim_plan *plan, const double *x, double *y);
int ibim_matvec_device(ibim_plan *plan, const double *d_x, double *d_y);

/* ---- fused matvecs over several plans --------------------------------------
 * No counterpart in the reference, which applies one molecule's operator at a
 * time (electric.cpp:294-388 per GMRES step, one molecule per process).  A
 * group holds k (1..16) plans of one device, Chebyshev order (p <= 6) and
 * kernel family; ibim_group_matvec_device then computes y_m = A_m x_m for
 * every member with ONE launch per FMM stage (charges, P2M, each M2M and L2L
 * level, M2L, near field, L2P, combination), the members' blocks one after
 * another, on the group's own streams: small solves that would each leave
 * most of the GPU idle share its launches.  mixed = 0: the FP64 matvec of
 * every member; mixed != 0: the mixed-precision matvec of the members'
 * precision (all the same; the FP16 form for p <= 5).  Each member's result
 * is the one ibim_matvec_device gives (the same kernels on the same data;
 * bit for bit where those are deterministic).  A call is ordered after the
 * work already queued on every member's stream, and work queued there later
 * runs after it.  The group must be destroyed before any member, and a
 * member's plan must not be used elsewhere while a group call runs. */
typedef struct ibim_group ibim_group;
int ibim_group_create(ibim_plan *const *plans, int k, int mixed, ibim_group **out);
int ibim_group_matvec_device(ibim_group *g, const double *const *d_x, double *const *d_y);
/* stage timers of the fused launches follow the first member's selection
   (ibim_plan_profile_stages) at creation */
int ibim_group_stage_report(ibim_group *g, ibim_stage_report *out, int reset);
void ibim_group_destroy(ibim_group *g);

/* one black-box FMM: potential[t] = sum_s K(src_s, tgt_t) charge[s] for one
 * of the 16 kernels (kernel_id = IBIM_K*), far field by Chebyshev FMM */
int ibim_kernel_fmm(ibim_plan *plan, int kernel_id, const double *charge, double *potential);
/* the same at m separate target points (AoS), as bbfmm::kernel with a target
 * set other than the sources (polarizeMap, electric.cpp:468-478) */
int ibim_kernel_fmm_targets(ibim_plan *plan, int kernel_id, int64_t m, const double *target_xyz,
                            const double *charge, double *potential);

/* ---- either side of the solve -------------------------------------------- */
/* right-hand side (length 2n) from m atoms (AoS centres, charges) */
int ibim_rhs(ibim_plan *plan, int64_t m, const double *atom_xyz, const double *atom_q, double *b);
int ibim_rhs_device(ibim_plan *plan, int64_t m, const double *d_atom_xyz, const double *d_atom_q,
                    double *d_b);

/* reaction potential at m target points (atom centres) from a solution x */
int ibim_polarize(ibim_plan *plan, int64_t m, const double *atom_xyz, const double *x, double *phi_reac);
int ibim_polarize_device(ibim_plan *plan, int64_t m, const double *d_atom_xyz, const double *d_x,
                         double *d_phi_reac);
/* ibim_polarize_device without any device allocation, for callers that
 * cannot allocate (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Its temporaries
 * (the sorted targets, their ranges in the tree, the reaction fields) live in
 * a 64 MiB array that each GPU holds from the library's load to the end of
 * the process; a call whose targets and tree would need more (about
 * 104 m + 27 boxes + 1 MiB bytes) is rejected with IBIM_ERR_ARG.  The far
 * field evaluates the V and X lists directly and the near field per target
 * leaf, so the result agrees with ibim_polarize_device to rounding (1e-10
 * relative), not bit for bit.  Calls on one GPU share the array and are
 * serialised by a lock inside the library (each returns when its result is
 * in d_phi_reac); the array keeps the last call's contents. */
int ibim_polarize_device_noalloc(ibim_plan *plan, int64_t m, const double *d_atom_xyz, const double *d_x,
                                 double *d_phi_reac);

/* restarted GMRES(restart) on A x = b with this plan's matvec; x holds the
 * initial guess on entry.  Returns IBIM_OK, IBIM_NOT_CONVERGED or an error. */
int ibim_gmres(ibim_plan *plan, const double *b, double *x, int restart, int maxit, double tol,
               ibim_solve_info *info);
int ibim_gmres_device(ibim_plan *plan, const double *d_b, double *d_x, int restart, int maxit,
                      double tol, ibim_solve_info *info);
/* on: ibim_gmres[_device] print the reference GMRES's table to stdout, one
 * line per iteration (iteration, relative residual, seconds), and the
 * CONVERGED / NOT CONVERGED banner (gmres.cpp:79-106, 135-162); the mixed
 * forms add one line per FP64 refinement step.  Off by default. */
int ibim_plan_set_trace(ibim_plan *plan, int on);

/* ---- inspection ----------------------------------------------------------- */
int ibim_tree_stats_get(ibim_plan *plan, ibim_tree_stats *out);
/* per-node dump (arrays of length n_nodes; centre/radius 3 per node) in the
 * plan's node order (breadth first, children of a node consecutive) */
int ibim_tree_export(ibim_plan *plan, int *level, int *parent, int *slot, int *leaf, int *empty,
                     int *nsrc, double *centre, double *radius, int *nu, int *nv, int *nw, int *nx);
/* interaction list (kind 0=U 1=V 2=W 3=X) of every node: offsets (n_nodes+1)
 * and items (offsets[n_nodes]) */
int ibim_tree_lists(ibim_plan *plan, int kind, int *offsets, int *items, int64_t cap);
/* source points of node id, original indices ascending */
int ibim_tree_points(ibim_plan *plan, int id, int *out, int64_t cap);

/* ---- geometry pre-step: level set and tube quadrature nodes ---------------
 * The solvent-excluded surface as a level set on a uniform grid and the IBIM
 * quadrature nodes in the tube |phi| < thickness*h around it.  Grid index
 * I = i*ny*nz + j*nz + k, coordinates lo + (i,j,k)*h (levelset.cpp:28-37, 554-558). */
typedef struct ibim_geom_params {
    double h;              /* grid spacing, in the atoms' length unit */
    double lo[3];          /* grid origin (sx, sy, sz) */
    int dims[3];           /* nx, ny, nz; all 0: derived from the atoms and `pad` */
    double pad;            /* margin beyond the probe-expanded atoms when dims are 0 */
    double probe;          /* probe radius (main.cpp:35: 1.4 Angstrom) */
    int band;              /* expand window margin in cells (main.cpp:45: 8) */
    int thickness;         /* tube half width in cells, lvlset_thickness */
    double thres;          /* reinitialisation gradient tolerance, lvlset_thres */
    double reinit_time;    /* final time of each reinitialisation, reinit_step */
    double scale;          /* s: nodes/s, weights*s*(h/s)^3 (electric.cpp:17-41); 1 for Angstrom grids */
    double init;           /* the grid's value before expand, <= 0 (main.cpp:43 starts from -2*size);
                              0: -2*max(dims)*h */
} ibim_geom_params;

typedef struct ibim_surface ibim_surface;

typedef struct ibim_surface_info {
    int64_t n_nodes;       /* tube points */
    int dims[3];
    double lo[3];
    double h;
    int64_t inclusions;    /* cavity grid points pushed out of the tube (levelset.cpp:526-537) */
    int evolve_steps, reinit_steps[2];
    double quality[2];     /* mean |grad phi| - 1 and bad fraction in the tube (levelset.cpp:398-435) */
    double area;           /* sum of weights */
} ibim_surface_info;

/* expand + evolve (probe) + flip + reinitialise twice + Surface, as the
 * reference pipeline main.cpp:47-62 (levelset.cpp:91-287, 437-586) */
int ibim_surface_from_atoms(const ibim_geom_params *gp, int64_t m, const double *atom_xyz, const double *radii,
                            ibim_surface **out);
/* Surface (levelset.cpp:437-586) of a level set given on the grid (host array) */
int ibim_surface_from_grid(const ibim_geom_params *gp, const double *phi, ibim_surface **out);
int ibim_surface_get_info(const ibim_surface *s, ibim_surface_info *info);
/* host copies (AoS xyz and normals, weights) */
int ibim_surface_copy(const ibim_surface *s, double *xyz, double *normals, double *weights);
/* device arrays owned by the surface */
int ibim_surface_device(const ibim_surface *s, const double **d_xyz, const double **d_normals,
                        const double **d_weights);
void ibim_surface_destroy(ibim_surface *s);

/* ---- molecules ---------------------------------------------------------------
 * Seeded synthetic globule: self-avoiding random walk of 3.8 A steps confined
 * to a sphere, radii U[1.2,2.0], charges U[-0.5,0.5] shifted to net -2 e.
 * Arrays are sized m (xyz 3m). */
int ibim_synthetic_globule(int64_t m, uint64_t seed, double confine_radius, double *xyz, double *radii,
                           double *charges);

/* ---- serving: continuous batching of PB solves -------------------------------
 * A request is one molecule: the level set and tube nodes are built, then
 * the plan, right-hand side, GMRES and reaction potential, and the result is
 * the solvation energy.  Up to `slots` requests run at once, each on its own
 * HIP streams with its state resident in HBM; a request is admitted when a
 * slot is free and its estimated device footprint fits next to the running
 * ones within `mem_budget` bytes (0: 90% of the free device memory at
 * creation).  A free slot starts the first queued request (in submission
 * order) that fits, so a small request may pass a large one that has to wait
 * for memory; a request passed over 8 times holds back all later ones until
 * it starts, and one that could never fit runs alone.  This replaces the
 * reference's one-molecule-per-process executable (main.cpp:14-79) for
 * streams of molecules. */
typedef struct ibim_server ibim_server;

typedef struct ibim_request {
    int64_t n_atoms;
    const double *atom_xyz;   /* 3 per atom; copied at submission */
    const double *radii;
    const double *charges;
    ibim_geom_params geom;    /* grid spacing etc.; dims 0 = fitted to the atoms */
    ibim_params fmm;          /* p, leaf, kappa, eps, vacant radius, precision */
    int restart, maxit;
    double tol;
} ibim_request;

typedef struct ibim_result {
    int status;               /* IBIM_OK, IBIM_NOT_CONVERGED or an error code */
    double energy;            /* (1/2) sum q phi_reac */
    int64_t n_points;         /* tube points */
    ibim_solve_info info;
    double t_submit, t_start, t_end;  /* seconds on the server's steady clock */
    int64_t footprint_est, footprint; /* device bytes: admission estimate, measured plan + surface */
} ibim_result;

int ibim_server_create(int slots, int64_t mem_budget, ibim_server **out);
/* copies the request; returns its id (>= 0) in *id */
int ibim_server_submit(ibim_server *s, const ibim_request *req, int64_t *id);
/* blocks until request `id` has finished */
int ibim_server_wait(ibim_server *s, int64_t id, ibim_result *out);
/* seconds on the server's clock (for arrival schedules) */
double ibim_server_now(ibim_server *s);
/* finishes every queued request, then stops the workers */
void ibim_server_destroy(ibim_server *s);
/* stage timers in the plan of every request started after the call (see
 * ibim_plan_profile_stages), and their reports summed over the requests that
 * have finished.  No counterpart in the reference. */
int ibim_server_profile_stages(ibim_server *s, unsigned stage_mask);
int ibim_server_stage_report(ibim_server *s, ibim_stage_report *out, int reset);
/* Fused stage launches across concurrent solves (no counterpart in the
 * reference).  Requests of at most max_points tube points run their GMRES
 * matvecs through a shared batcher: a solve that needs a matvec waits up to
 * wait_us microseconds for the other such solves to need one too, and then
 * all waiting matvecs of one form run as one fused group matvec (see
 * ibim_group_create); larger requests, and solves that find no partner,
 * run their own.  Each result is the same as alone (same kernels, same
 * data).  max_points = 0 switches fusion off (the default); applies to
 * requests started after the call. */
int ibim_server_set_fusion(ibim_server *s, int64_t max_points, int wait_us);
/* fused launches so far: group matvecs and the member matvecs they ran */
int ibim_server_fusion_stats(ibim_server *s, int64_t *group_matvecs, int64_t *member_matvecs);

const char *ibim_last_error(void);
const char *ibim_version(void);

#ifdef __cplusplus
}
#endif
#endif
