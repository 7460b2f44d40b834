/*
 * ibim_levelset_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's geometry pre-step
 * (GaZ3ll3/ibim-levelset levelset.cpp and main.cpp:47-62), used by tests/ to
 * check the GPU level set and tube (csrc/ibim_geometry.hip).  Nothing in the
 * product links or calls this file.
 *
 *   grid access      Grid::get clamps, Grid::set does not   levelset.cpp:28-37
 *   expand           levelset.cpp:91-109
 *   evolve           levelset.cpp:111-187
 *   reinitialize     levelset.cpp:191-292
 *   getNorm          levelset.cpp:294-301
 *   setWindow        levelset.cpp:309-327
 *   setGradient      levelset.cpp:330-396 (WENO5, EPS_LS = 1e-6)
 *   countGradient    levelset.cpp:398-435
 *   Surface          levelset.cpp:437-586 (flood fill, inclusions, nodes)
 *   unit conversion  electric.cpp:17-41 (nodes / s, weights * s * (h / s)^3)
 *
 * One deviation, shared with the GPU code and stated in the design notes:
 * expand skips window cells outside the grid, where the reference writes
 * out of bounds.  Each stage writes every cell once from the previous
 * stage's grid, so the OpenMP loops (as the reference's) give the same
 * values in any order; only countGradient's mean error is summed in a
 * thread-dependent order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ibim_oracle.h"

#define SQR(x) ((x) * (x))
#define LS_PI 3.14159265358979323846

typedef struct {
    int nx, ny, nz;
    double sx, sy, sz, dx;
} lsgrid;

static size_t gidx(const lsgrid *d, int i, int j, int k) { return ((size_t)i * d->ny + j) * d->nz + k; }

/* Grid::get, levelset.cpp:28-33: indices clamped to the grid */
static double gget(const double *g, const lsgrid *d, int i, int j, int k) {
    i = i < 0 ? 0 : (i > d->nx - 1 ? d->nx - 1 : i);
    j = j < 0 ? 0 : (j > d->ny - 1 ? d->ny - 1 : j);
    k = k < 0 ? 0 : (k > d->nz - 1 ? d->nz - 1 : k);
    return g[gidx(d, i, j, k)];
}

/* setWindow (levelset.cpp:309-327): the 7 values along each axis around (i, j, k) */
static void set_window(const double *g, const lsgrid *d, int i, int j, int k, double w[21]) {
    for (int l = 0; l < 7; ++l) {
        w[l] = gget(g, d, i - 3 + l, j, k);
        w[7 + l] = gget(g, d, i, j - 3 + l, k);
        w[14 + l] = gget(g, d, i, j, k - 3 + l);
    }
}

/* setGradient (levelset.cpp:330-396): WENO5 forward (uxp) and backward
 * (uxn) one-sided derivatives along dir */
static void set_gradient(int dir, const double *window, double dx, double uxp[3], double uxn[3]) {
    const double EPS_LS = 1e-6;
    const double *w = window + 7 * dir;
    double a = w[1] - w[0], b = w[2] - w[1], c = w[3] - w[2], d = w[4] - w[3], e = w[5] - w[4], f = w[6] - w[5];
    const double df = (-b + 7.0 * (c + d) - e) / 12.0;
    a = b - a;
    b = c - b;
    c = d - c;
    d = e - d;
    e = f - e;
    double is0 = 13.0 * SQR(e - d) + 3.0 * SQR(e - 3.0 * d);
    double is1 = 13.0 * SQR(d - c) + 3.0 * SQR(d + c);
    double is2 = 13.0 * SQR(c - b) + 3.0 * SQR(3.0 * c - b);
    double a0 = 1.0 / SQR(EPS_LS + is0), a1 = 6.0 / SQR(EPS_LS + is1), a2 = 3.0 / SQR(EPS_LS + is2);
    double w1 = 1.0 / (a0 + a1 + a2);
    double w0 = a0 * w1, w2 = a2 * w1;
    uxp[dir] = (df + w0 * (e - 2 * d + c) / 3.0 + (w2 - 0.5) * (d - 2.0 * c + b) / 6.0) / dx;
    is0 = 13.0 * SQR(a - b) + 3.0 * SQR(a - 3.0 * b);
    is1 = 13.0 * SQR(b - c) + 3.0 * SQR(b + c);
    is2 = 13.0 * SQR(c - d) + 3.0 * SQR(3.0 * c - d);
    a0 = 1.0 / SQR(EPS_LS + is0);
    a1 = 6.0 / SQR(EPS_LS + is1);
    a2 = 3.0 / SQR(EPS_LS + is2);
    w1 = 1.0 / (a0 + a1 + a2);
    w0 = a0 * w1;
    w2 = a2 * w1;
    uxn[dir] = (df - w0 * (a - 2 * b + c) / 3.0 - (w2 - 0.5) * (b - 2 * c + d) / 6.0) / dx;
}

/* getNorm (levelset.cpp:294-301): sqrt(sum max(A+, (-B)+)^2) */
static double get_norm(const double A[3], const double B[3]) {
    double v = 0.0;
    for (int q = 0; q < 3; ++q) {
        const double t = fmax(A[q] > 0. ? A[q] : 0., B[q] < 0. ? -B[q] : 0.);
        v += SQR(t);
    }
    return sqrt(v);
}

static void gradients(const double *g, const lsgrid *d, int i, int j, int k, double Dup[3], double Dun[3]) {
    double w[21];
    set_window(g, d, i, j, k, w);
    for (int dir = 0; dir < 3; ++dir) set_gradient(dir, w, d->dx, Dup, Dun);
}

/* countGradient (levelset.cpp:398-435): tube points and those whose
 * gradient norm is off by more than thres; the mean error */
static int64_t count_gradient(const double *g, const lsgrid *d, double thickness, double thres, int64_t *total,
                              double *mean_err) {
    int64_t tot = 0, bad = 0;
    double acc = 0.0;
#pragma omp parallel for collapse(3) reduction(+ : tot, bad, acc) schedule(static)
    for (int i = 0; i < d->nx; ++i)
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k)
                if (fabs(g[gidx(d, i, j, k)]) < thickness * d->dx) {
                    double Dup[3], Dun[3];
                    tot++;
                    gradients(g, d, i, j, k, Dup, Dun);
                    const double nr = get_norm(Dun, Dup);
                    if (fabs(nr - 1.0) > thres) bad++;
                    acc += fabs(nr - 1.0);
                }
    if (total) *total = tot;
    if (mean_err) *mean_err = tot ? acc / (double)tot : 0.0;
    return bad;
}

/* one step of the TVD-RK3 schemes of evolve (mode 0, levelset.cpp:123-183)
 * and reinitialize (mode 1, levelset.cpp:214-280) */
static double rk_rhs(int mode, const double *u, const double *phi0, const lsgrid *d, int i, int j, int k,
                     double dt, double vel, double eps) {
    double Dup[3], Dun[3];
    gradients(u, d, i, j, k, Dup, Dun);
    if (mode == 0) return dt * vel * get_norm(Dun, Dup);
    const double p0 = phi0[gidx(d, i, j, k)];
    const double sign = p0 / sqrt(SQR(p0) + eps);
    const double normDu = sign > 0. ? get_norm(Dun, Dup) : get_norm(Dup, Dun);
    return dt * vel * sign * (normDu - 1.0);
}

static void rk_step(int mode, double *g, const double *phi0, double *u1, double *u2, const lsgrid *d, double dt,
                    double vel, double eps) {
#pragma omp parallel for collapse(3) schedule(static)
    for (int i = 0; i < d->nx; ++i)
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k) {
                const size_t I = gidx(d, i, j, k);
                u1[I] = g[I] - rk_rhs(mode, g, phi0, d, i, j, k, dt, vel, eps);
            }
#pragma omp parallel for collapse(3) schedule(static)
    for (int i = 0; i < d->nx; ++i)
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k) {
                const size_t I = gidx(d, i, j, k);
                u2[I] = (3 * g[I] + u1[I] - rk_rhs(mode, u1, phi0, d, i, j, k, dt, vel, eps)) / 4.0;
            }
    /* the third stage reads u2 only, so g can take its result in place */
#pragma omp parallel for collapse(3) schedule(static)
    for (int i = 0; i < d->nx; ++i)
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k) {
                const size_t I = gidx(d, i, j, k);
                g[I] = (g[I] + 2 * (u2[I] - rk_rhs(mode, u2, phi0, d, i, j, k, dt, vel, eps))) / 3.0;
            }
}

/* evolve (mode 0) or reinitialize (mode 1) to final_t; returns the steps taken */
static int rk_run(int mode, double *g, const double *phi0, const lsgrid *d, double final_t, double vel,
                  double cfl, double thickness, double thres, double quality[2]) {
    double dt = cfl * d->dx / vel;
    const double eps = d->dx * d->dx;
    const int num_steps = (int)(final_t / dt) + 1;
    dt = final_t / (double)num_steps;
    const size_t n = (size_t)d->nx * d->ny * d->nz;
    double *u1 = (double *)malloc(n * sizeof(double)), *u2 = (double *)malloc(n * sizeof(double));
    int step = 0;
    if (mode == 0) {
        for (step = 0; step < num_steps; ++step) rk_step(0, g, phi0, u1, u2, d, dt, vel, eps);
    } else {
        int64_t indices = count_gradient(g, d, thickness, thres, NULL, NULL);
        while (indices != 0 && step < num_steps) {
            step++;
            rk_step(1, g, phi0, u1, u2, d, dt, vel, eps);
            indices = count_gradient(g, d, thickness, thres, NULL, NULL);
        }
        if (quality) {
            int64_t tot = 0;
            const int64_t bad = count_gradient(g, d, thickness, thres, &tot, &quality[0]);
            quality[1] = tot ? (double)bad / (double)tot : 0.0;
        }
    }
    free(u1);
    free(u2);
    return step;
}

/* Surface (levelset.cpp:437-586) of g, in grid order, converted to the
 * atoms' unit (electric.cpp:17-41); returns the node count (the arrays, if
 * given, must hold cap nodes) or -1 */
static int64_t surface(double *g, const lsgrid *d, int thickness, double scale, double *xyz, double *nrm,
                       double *w, int64_t cap, int64_t *inclusions) {
    const double tube = thickness * d->dx;
    const size_t n = (size_t)d->nx * d->ny * d->nz;
    unsigned char *visited = (unsigned char *)calloc(n, 1);
    int *queue = (int *)malloc(n * sizeof(int));
    size_t qh = 0, qt = 0;
    queue[qt++] = 0;
    visited[0] = 1;
    const int nyz = d->ny * d->nz;
    while (qh < qt) { /* breadth first from grid point 0 (levelset.cpp:446-524) */
        const int cur = queue[qh++];
        const int I = cur / nyz, J = cur / d->nz - I * d->ny, K = cur % d->nz;
        const int nb[6][3] = {{I + 1, J, K}, {I - 1, J, K}, {I, J + 1, K}, {I, J - 1, K}, {I, J, K + 1}, {I, J, K - 1}};
        for (int q = 0; q < 6; ++q) {
            const int ci = nb[q][0], cj = nb[q][1], ck = nb[q][2];
            if (ci < 0 || cj < 0 || ck < 0 || ci > d->nx - 1 || cj > d->ny - 1 || ck > d->nz - 1) continue;
            const int next = ci * nyz + cj * d->nz + ck;
            if (!visited[next] && g[cur] > -tube) {
                visited[next] = 1;
                queue[qt++] = next;
            }
        }
    }
    int64_t cnt = 0; /* inclusions (levelset.cpp:529-537) */
    const double inclusion_value = -(tube + 1e-15);
    for (size_t id = 0; id < n; ++id)
        if (!visited[id] && fabs(g[id]) < tube) {
            g[id] = inclusion_value;
            cnt++;
        }
    if (inclusions) *inclusions = cnt;
    free(visited);
    free(queue);
    const double s = scale > 0 ? scale : 1.0, hs = d->dx / s;
    int64_t m = 0;
    for (int i = 0; i < d->nx; ++i) /* levelset.cpp:540-583 */
        for (int j = 0; j < d->ny; ++j)
            for (int k = 0; k < d->nz; ++k) {
                const double dist = gget(g, d, i, j, k);
                if (!(fabs(dist) < tube)) continue;
                if (xyz) {
                    if (m >= cap) return -1;
                    double Dup[3], Dun[3];
                    gradients(g, d, i, j, k, Dup, Dun);
                    const double P[3] = {d->sx + i * d->dx, d->sy + j * d->dx, d->sz + k * d->dx};
                    for (int q = 0; q < 3; ++q) {
                        const double gr = (Dun[q] + Dup[q]) * 0.5;
                        xyz[3 * m + q] = (P[q] - gr * dist) / s;
                        nrm[3 * m + q] = gr;
                    }
                    w[m] = 0.5 * (1.0 + cos(LS_PI * dist / tube)) / tube * (s * hs * hs * hs);
                }
                m++;
            }
    return m;
}

/* the WENO5 one-sided derivatives at grid point (i, j, k) of g */
void orc_ls_gradient(const orc_lsparams *p, const double *g, int i, int j, int k, double Dup[3], double Dun[3]) {
    const lsgrid d = {p->dims[0], p->dims[1], p->dims[2], p->lo[0], p->lo[1], p->lo[2], p->h};
    gradients(g, &d, i, j, k, Dup, Dun);
}

double orc_ls_norm(const double A[3], const double B[3]) { return get_norm(A, B); }

/* evolve (mode 0) or reinitialize (mode 1, phi0 required) g in place */
int orc_ls_run(const orc_lsparams *p, int mode, double *g, const double *phi0, double final_t, double vel, double cfl,
               double quality[2]) {
    const lsgrid d = {p->dims[0], p->dims[1], p->dims[2], p->lo[0], p->lo[1], p->lo[2], p->h};
    return rk_run(mode, g, phi0, &d, final_t, vel, cfl, p->thickness, p->thres, quality);
}

int64_t orc_surface_from_grid(const orc_lsparams *p, const double *phi, double *xyz, double *nrm, double *w,
                              int64_t cap, int64_t *inclusions) {
    const lsgrid d = {p->dims[0], p->dims[1], p->dims[2], p->lo[0], p->lo[1], p->lo[2], p->h};
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    double *g = (double *)malloc(n * sizeof(double));
    memcpy(g, phi, n * sizeof(double));
    const int64_t m = surface(g, &d, p->thickness, p->scale, xyz, nrm, w, cap, inclusions);
    free(g);
    return m;
}

int64_t orc_surface_from_atoms(const orc_lsparams *p, int m, const double *axyz, const double *radii,
                               double init, double *xyz, double *nrm, double *w, int64_t cap, orc_lsinfo *info,
                               double *g_out) {
    const lsgrid d = {p->dims[0], p->dims[1], p->dims[2], p->lo[0], p->lo[1], p->lo[2], p->h};
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    double *g = (double *)malloc(n * sizeof(double)), *phi0 = (double *)malloc(n * sizeof(double));
    for (size_t I = 0; I < n; ++I) g[I] = init; /* the initial grid (main.cpp:43) */
    for (int a = 0; a < m; ++a) {              /* expand (levelset.cpp:91-109) */
        const double rp = radii[a] + p->probe;
        const int r = (int)(rp / d.dx) + p->band;
        const int icx = (int)((axyz[3 * a] - d.sx) / d.dx), icy = (int)((axyz[3 * a + 1] - d.sy) / d.dx),
                  icz = (int)((axyz[3 * a + 2] - d.sz) / d.dx);
        for (int ia = icx - r; ia <= icx + r; ++ia)
            for (int ib = icy - r; ib <= icy + r; ++ib)
                for (int ic = icz - r; ic <= icz + r; ++ic) {
                    if (ia < 0 || ib < 0 || ic < 0 || ia >= d.nx || ib >= d.ny || ic >= d.nz) continue;
                    const double px = d.sx + ia * d.dx - axyz[3 * a], py = d.sy + ib * d.dx - axyz[3 * a + 1],
                                 pz = d.sz + ic * d.dx - axyz[3 * a + 2];
                    double dist = rp - sqrt(px * px + py * py + pz * pz);
                    const size_t I = gidx(&d, ia, ib, ic);
                    if (g[I] > dist) dist = g[I];
                    g[I] = dist;
                }
    }
    orc_lsinfo inf;
    memset(&inf, 0, sizeof(inf));
    /* evolve inward by the probe radius (main.cpp:51), flip (main.cpp:53),
     * reinitialise twice from the current level set (main.cpp:54-60) */
    inf.evolve_steps = rk_run(0, g, NULL, &d, p->probe, 1.0, 0.2, p->thickness, p->thres, NULL);
    for (size_t I = 0; I < n; ++I) g[I] *= -1.0;
    for (int pass = 0; pass < 2; ++pass) {
        memcpy(phi0, g, n * sizeof(double));
        inf.reinit_steps[pass] = rk_run(1, g, phi0, &d, p->reinit_time, 1.0, 0.5, p->thickness, p->thres, inf.quality);
    }
    if (g_out) memcpy(g_out, g, n * sizeof(double));
    const int64_t cnt = surface(g, &d, p->thickness, p->scale, xyz, nrm, w, cap, &inf.inclusions);
    inf.n_nodes = cnt;
    if (info) *info = inf;
    free(g);
    free(phi0);
    return cnt;
}
