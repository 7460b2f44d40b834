/*
 * ibim_oracle.c -- CPU restatement of the reference FMM matvec path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the
 * CPU baseline timed beside it.  The product never links or calls it.
 * See ibim_oracle.h for what each function restates and how it is pinned.
 *
 * The structure deliberately follows the reference algorithm rather than the
 * GPU design: one tree per kernel, dense r x r transfer matrices R[8], kernel
 * matrices evaluated entry by entry and applied as matrix-vector products.
 * Parallelism is OpenMP over the boxes of a level (the reference uses OpenMP
 * task recursion over the same boxes, bbfmm.h:245-269,522-529,614-621).
 */
#include "ibim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_EPS 1e-12               /* utils.h:40 */
#define ORC_PI 3.14159265358979323846
#define SQR(x) ((x) * (x))

/* ------------------------------------------------------------------------ */
/* small growable int list with set semantics (unordered_set in bbfmm.h:78) */

static void list_insert(int **a, int *n, int *cap, int v) {
    for (int i = 0; i < *n; ++i)
        if ((*a)[i] == v) return;
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 8;
        *a = (int *)realloc(*a, (size_t)(*cap) * sizeof(int));
    }
    (*a)[(*n)++] = v;
}

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------ */
/* tree construction: bbfmm.cpp:8-138                                        */

static int new_node(orc_tree *t, int level, int slot) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = (orc_node *)realloc(t->nodes, (size_t)t->cap * sizeof(orc_node));
    }
    orc_node *n = &t->nodes[t->nnodes];
    memset(n, 0, sizeof(*n));
    n->parent = -1;
    for (int i = 0; i < 8; ++i) n->child[i] = -1;
    n->level = level;
    n->slot = slot;
    return t->nnodes++;
}

/* bbfmm.cpp:43-65: root box = bounding box of the SOURCES, per-axis radius */
static void center_radius(const double *src, int n, double *c, double *r) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = src[d];
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double v = src[3 * i + d];
            if (v > hi[d]) hi[d] = v;
            if (v < lo[d]) lo[d] = v;
        }
    for (int d = 0; d < 3; ++d) {
        c[d] = (hi[d] + lo[d]) / 2.0;
        r[d] = (hi[d] - lo[d]) / 2.0;
    }
}

/* child slot of a point: 4*z + 2*y + x, each bit `coord < centre ? 0 : 1`
 * (bbfmm.cpp:106-111) */
static int child_slot(const double *pt, const double *c) {
    int zb = pt[2] < c[2] ? 0 : 1;
    int yb = pt[1] < c[1] ? 0 : 1;
    int xb = pt[0] < c[0] ? 0 : 1;
    return 4 * zb + 2 * yb + xb;
}

/* bbfmm.cpp:67-138 (recursive, children created eight at a time and
 * numbered consecutively, depth first) */
static void assign_children(orc_tree *t, int id) {
    orc_node *n = &t->nodes[id];
    if (n->nsrc == 0) {
        n->leaf = 1;
        n->empty = 1;
        return;
    }
    if (n->nsrc <= t->max_point || n->level == t->max_level) {
        n->leaf = 1;
        if (t->depth < n->level) t->depth = n->level;
        return;
    }
    int lvl = n->level;
    for (int i = 0; i < 8; ++i) {
        int cid = new_node(t, lvl + 1, i);
        n = &t->nodes[id]; /* storage may have moved */
        orc_node *c = &t->nodes[cid];
        n->child[i] = cid;
        c->parent = id;
        c->c[0] = n->c[0] + ((i & 1) - 0.5) * n->r[0];
        c->c[1] = n->c[1] + (((i >> 1) & 1) - 0.5) * n->r[1];
        c->c[2] = n->c[2] + ((i >> 2) - 0.5) * n->r[2];
        c->r[0] = n->r[0] * 0.5;
        c->r[1] = n->r[1] * 0.5;
        c->r[2] = n->r[2] * 0.5;
    }
    /* distribute sources then targets, keeping ascending index order */
    for (int pass = 0; pass < 2; ++pass) {
        int cnt[8] = {0};
        int np = pass == 0 ? n->nsrc : n->ntgt;
        int *idx = pass == 0 ? n->src : n->tgt;
        const double *pts = pass == 0 ? t->src : t->tgt;
        int *slot = (int *)malloc((size_t)(np > 0 ? np : 1) * sizeof(int));
        for (int k = 0; k < np; ++k) {
            slot[k] = child_slot(pts + 3 * (size_t)idx[k], n->c);
            cnt[slot[k]]++;
        }
        int *lists[8];
        for (int i = 0; i < 8; ++i) lists[i] = (int *)malloc((size_t)(cnt[i] > 0 ? cnt[i] : 1) * sizeof(int));
        int fill[8] = {0};
        for (int k = 0; k < np; ++k) lists[slot[k]][fill[slot[k]]++] = idx[k];
        for (int i = 0; i < 8; ++i) {
            orc_node *c = &t->nodes[n->child[i]];
            if (pass == 0) {
                c->src = lists[i];
                c->nsrc = cnt[i];
            } else {
                c->tgt = lists[i];
                c->ntgt = cnt[i];
            }
        }
        free(slot);
    }
    for (int i = 0; i < 8; ++i) assign_children(t, t->nodes[id].child[i]);
}

/* bbfmm.cpp:244-258 */
static int find_node(const orc_tree *t, const double *p) {
    int id = 0;
    for (;;) {
        const orc_node *n = &t->nodes[id];
        if (fabs(n->c[0] - p[0]) < ORC_EPS && fabs(n->c[1] - p[1]) < ORC_EPS &&
            fabs(n->c[2] - p[2]) < ORC_EPS)
            return id;
        if (n->leaf) return id;
        int xb = n->c[0] > p[0] ? 0 : 1;
        int yb = n->c[1] > p[1] ? 0 : 1;
        int zb = n->c[2] > p[2] ? 0 : 1;
        id = n->child[4 * zb + 2 * yb + xb];
    }
}

/* bbfmm.cpp:260-280: boxes touch on a face, edge or corner */
static int is_adjacent(const orc_tree *t, int a, int b) {
    const orc_node *A = &t->nodes[a], *B = &t->nodes[b];
    double dx = fabs(A->c[0] - B->c[0]), dy = fabs(A->c[1] - B->c[1]), dz = fabs(A->c[2] - B->c[2]);
    double rx = fabs(A->r[0] + B->r[0]), ry = fabs(A->r[1] + B->r[1]), rz = fabs(A->r[2] + B->r[2]);
    int rdx = rx >= dx - ORC_EPS, rdy = ry >= dy - ORC_EPS, rdz = rz >= dz - ORC_EPS;
    int xa = (fabs(dx - rx) < ORC_EPS) && (rdy && rdz);
    int ya = (fabs(dy - ry) < ORC_EPS) && (rdx && rdz);
    int za = (fabs(dz - rz) < ORC_EPS) && (rdy && rdx);
    return xa || ya || za;
}

/* bbfmm.cpp:158-241 */
static void build_node(orc_tree *t, int id, const double *min_p, const double *max_p) {
    orc_node *n = &t->nodes[id];
    if (n->parent != -1) {
        const orc_node *pn = &t->nodes[n->parent];
        double dx = n->r[0], dy = n->r[1], dz = n->r[2];
        double xs = pn->c[0] - dx, ys = pn->c[1] - dy, zs = pn->c[2] - dz;
        int qcap = 64, *queue = (int *)malloc((size_t)qcap * sizeof(int));
        for (int xi = -2; xi < 4; ++xi)
            for (int yi = -2; yi < 4; ++yi)
                for (int zi = -2; zi < 4; ++zi) {
                    double cur[3] = {xs + 2 * xi * dx, ys + 2 * yi * dy, zs + 2 * zi * dz};
                    int inside = cur[0] <= max_p[0] + ORC_EPS && cur[1] <= max_p[1] + ORC_EPS &&
                                 cur[2] <= max_p[2] + ORC_EPS && cur[0] >= min_p[0] - ORC_EPS &&
                                 cur[1] >= min_p[1] - ORC_EPS && cur[2] >= min_p[2] - ORC_EPS;
                    int self = fabs(cur[0] - n->c[0]) < ORC_EPS && fabs(cur[1] - n->c[1]) < ORC_EPS &&
                               fabs(cur[2] - n->c[2]) < ORC_EPS;
                    if (!inside || self) continue;
                    int cid = find_node(t, cur);
                    int adj = is_adjacent(t, id, cid);
                    const orc_node *cn = &t->nodes[cid];
                    if (cn->level < n->level) {
                        if (adj) {
                            if (cn->leaf) list_insert(&n->u, &n->nu, &n->cu, cid);
                        } else {
                            list_insert(&n->x, &n->nx, &n->cx, cid);
                        }
                    }
                    if (cn->level == n->level) {
                        if (!adj) {
                            list_insert(&n->v, &n->nv, &n->cv, cid);
                        } else if (n->leaf) {
                            int head = 0, tail = 0;
                            queue[tail++] = cid;
                            while (head < tail) {
                                int f = queue[head++];
                                const orc_node *fn = &t->nodes[f];
                                if (!is_adjacent(t, f, id)) {
                                    list_insert(&n->w, &n->nw, &n->cw, f);
                                } else if (fn->leaf) {
                                    list_insert(&n->u, &n->nu, &n->cu, f);
                                } else {
                                    if (tail + 8 > qcap) {
                                        /* compact consumed entries, then grow */
                                        memmove(queue, queue + head, (size_t)(tail - head) * sizeof(int));
                                        tail -= head;
                                        head = 0;
                                        if (tail + 8 > qcap) {
                                            qcap *= 2;
                                            queue = (int *)realloc(queue, (size_t)qcap * sizeof(int));
                                        }
                                    }
                                    for (int c = 0; c < 8; ++c) queue[tail++] = fn->child[c];
                                }
                            }
                        }
                    }
                }
        free(queue);
    }
    if (n->leaf) list_insert(&n->u, &n->nu, &n->cu, id);
    /* deterministic order (the reference iterates an unordered_set) */
    qsort(n->u, (size_t)n->nu, sizeof(int), cmp_int);
    qsort(n->v, (size_t)n->nv, sizeof(int), cmp_int);
    qsort(n->w, (size_t)n->nw, sizeof(int), cmp_int);
    qsort(n->x, (size_t)n->nx, sizeof(int), cmp_int);
}

int orc_tree_build(orc_tree *t, const double *src, int nsrc, const double *tgt, int ntgt,
                   int max_point, int max_level) {
    memset(t, 0, sizeof(*t));
    if (nsrc <= 0) return -1; /* bbfmm.cpp:44 asserts a non-empty source set */
    t->nsrc = nsrc;
    t->ntgt = ntgt;
    t->src = (double *)malloc(3 * (size_t)nsrc * sizeof(double));
    memcpy(t->src, src, 3 * (size_t)nsrc * sizeof(double));
    t->tgt = (double *)malloc(3 * (size_t)(ntgt > 0 ? ntgt : 1) * sizeof(double));
    if (ntgt > 0) memcpy(t->tgt, tgt, 3 * (size_t)ntgt * sizeof(double));
    t->max_point = max_point;
    t->max_level = max_level;
    t->depth = 0;
    int root = new_node(t, 0, 0);
    orc_node *r = &t->nodes[root];
    center_radius(src, nsrc, r->c, r->r);
    r->nsrc = nsrc;
    r->ntgt = ntgt;
    r->src = (int *)malloc((size_t)nsrc * sizeof(int));
    r->tgt = (int *)malloc((size_t)(ntgt > 0 ? ntgt : 1) * sizeof(int));
    for (int i = 0; i < nsrc; ++i) r->src[i] = i;
    for (int i = 0; i < ntgt; ++i) r->tgt[i] = i;
    assign_children(t, root);

    const orc_node *rt = &t->nodes[0];
    double min_p[3], max_p[3];
    for (int d = 0; d < 3; ++d) {
        min_p[d] = rt->c[d] - rt->r[d];
        max_p[d] = rt->c[d] + rt->r[d];
    }
    int nn = t->nnodes;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nn; ++i) build_node(t, i, min_p, max_p);
    return 0;
}

void orc_tree_free(orc_tree *t) {
    for (int i = 0; i < t->nnodes; ++i) {
        orc_node *n = &t->nodes[i];
        free(n->src);
        free(n->tgt);
        free(n->u);
        free(n->v);
        free(n->w);
        free(n->x);
    }
    free(t->nodes);
    free(t->src);
    free(t->tgt);
    memset(t, 0, sizeof(*t));
}

int orc_tree_nnodes(const orc_tree *t) { return t->nnodes; }

void orc_tree_export(const orc_tree *t, int *level, int *parent, int *slot, int *leaf, int *empty,
                     int *nsrc, int *ntgt, double *centre, double *radius, int *nu, int *nv, int *nw,
                     int *nx) {
    for (int i = 0; i < t->nnodes; ++i) {
        const orc_node *n = &t->nodes[i];
        level[i] = n->level;
        parent[i] = n->parent;
        slot[i] = n->slot;
        leaf[i] = n->leaf;
        empty[i] = n->empty;
        nsrc[i] = n->nsrc;
        ntgt[i] = n->ntgt;
        for (int d = 0; d < 3; ++d) {
            centre[3 * i + d] = n->c[d];
            radius[3 * i + d] = n->r[d];
        }
        nu[i] = n->nu;
        nv[i] = n->nv;
        nw[i] = n->nw;
        nx[i] = n->nx;
    }
}

int orc_tree_list(const orc_tree *t, int id, int kind, int *out, int cap) {
    const orc_node *n = &t->nodes[id];
    const int *a = kind == 0 ? n->u : kind == 1 ? n->v : kind == 2 ? n->w : n->x;
    int len = kind == 0 ? n->nu : kind == 1 ? n->nv : kind == 2 ? n->nw : n->nx;
    for (int i = 0; i < len && i < cap; ++i) out[i] = a[i];
    return len;
}

int orc_tree_points(const orc_tree *t, int id, int which, int *out, int cap) {
    const orc_node *n = &t->nodes[id];
    const int *a = which == 0 ? n->src : n->tgt;
    int len = which == 0 ? n->nsrc : n->ntgt;
    for (int i = 0; i < len && i < cap; ++i) out[i] = a[i];
    return len;
}

/* ------------------------------------------------------------------------ */
/* Chebyshev operators: bbfmm.h:273-433                                     */

void orc_cheb_nodes(int p, double *x) {
    for (int i = 0; i < p; ++i) x[i] = -cos((i + 0.5) * ORC_PI / p);
}

/* T(i,k) = T_k(x_i), col-major n x p (bbfmm.h:280-303) */
void orc_cheb_poly(int p, int n, const double *x, double *T) {
    for (int i = 0; i < n; ++i) T[i] = 1.0;
    if (p > 1) {
        for (int i = 0; i < n; ++i) T[n + i] = x[i];
        for (int k = 2; k < p; ++k)
            for (int i = 0; i < n; ++i)
                T[(size_t)k * n + i] = 2.0 * x[i] * T[(size_t)(k - 1) * n + i] - T[(size_t)(k - 2) * n + i];
    }
}

/* S(i,m) = S_p(x_i, xbar_m) = (1/p)(2 sum_k T_k(x_i) T_k(xbar_m) - 1),
 * col-major n x p (bbfmm.h:305-326, 408-419) */
void orc_sp(int p, int n, const double *x, double *S) {
    double *xb = (double *)malloc((size_t)p * sizeof(double));
    double *Tb = (double *)malloc((size_t)p * p * sizeof(double));
    double *Tx = (double *)malloc((size_t)p * n * sizeof(double));
    orc_cheb_nodes(p, xb);
    orc_cheb_poly(p, p, xb, Tb);
    orc_cheb_poly(p, n, x, Tx);
    double inv = 1.0 / p;
    for (int m = 0; m < p; ++m)
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = 0; k < p; ++k) acc += Tx[(size_t)k * n + i] * Tb[(size_t)k * p + m];
            S[(size_t)m * n + i] = (2.0 * acc - 1.0) * inv;
        }
    free(xb);
    free(Tb);
    free(Tx);
}

/* R[id](i p^2 + j p + k, l p^2 + m p + n) = Tz(i,l) Ty(j,m) Tx(k,n) with
 * T[b] the p x p parent-to-child interpolation of child half b
 * (bbfmm.h:328-367). R is 8 consecutive col-major r x r blocks. */
void orc_transfer_R(int p, double *R) {
    int r = p * p * p;
    double *xb = (double *)malloc((size_t)p * sizeof(double));
    double *cn = (double *)malloc(2 * (size_t)p * sizeof(double));
    double *S = (double *)malloc(2 * (size_t)p * p * sizeof(double));
    orc_cheb_nodes(p, xb);
    for (int i = 0; i < p; ++i) {
        cn[i] = 0.5 * xb[i] + (-0.5);
        cn[p + i] = 0.5 * xb[i] + 0.5;
    }
    orc_sp(p, 2 * p, cn, S); /* (2p) x p col-major */
#define TR(b, i, l) S[(size_t)(l) * (2 * p) + (b) * p + (i)]
    for (int id = 0; id < 8; ++id) {
        int bx = id & 1, by = (id >> 1) & 1, bz = (id >> 2) & 1;
        double *Rb = R + (size_t)id * r * r;
        for (int i = 0; i < p; ++i)
            for (int j = 0; j < p; ++j)
                for (int k = 0; k < p; ++k)
                    for (int l = 0; l < p; ++l)
                        for (int m = 0; m < p; ++m)
                            for (int nn = 0; nn < p; ++nn) {
                                int row = i * p * p + j * p + k, col = l * p * p + m * p + nn;
                                Rb[(size_t)col * r + row] = TR(bz, i, l) * TR(by, j, m) * TR(bx, k, nn);
                            }
    }
#undef TR
    free(xb);
    free(cn);
    free(S);
}

/* leaf interpolation matrix (bbfmm.h:392-433): n x r col-major,
 * S(k, l p^2 + j p + i) = Sx(k,i) Sy(k,j) Sz(k,l) at the standardized
 * coordinates (x - c)/r of the listed points. */
static void leaf_matrix(int p, const double *pts, const int *idx, int n, const double *c,
                        const double *rad, double *out) {
    int r = p * p * p;
    double *st = (double *)malloc(3 * (size_t)(n > 0 ? n : 1) * sizeof(double));
    double *S[3];
    for (int d = 0; d < 3; ++d) {
        for (int k = 0; k < n; ++k) st[(size_t)d * n + k] = (pts[3 * (size_t)idx[k] + d] - c[d]) / rad[d];
        S[d] = (double *)malloc((size_t)p * (n > 0 ? n : 1) * sizeof(double));
        orc_sp(p, n, st + (size_t)d * n, S[d]);
    }
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < p; ++i)
            for (int j = 0; j < p; ++j)
                for (int l = 0; l < p; ++l)
                    out[(size_t)(l * p * p + j * p + i) * n + k] =
                        S[0][(size_t)i * n + k] * S[1][(size_t)j * n + k] * S[2][(size_t)l * n + k];
    (void)r;
    for (int d = 0; d < 3; ++d) free(S[d]);
    free(st);
}

/* ------------------------------------------------------------------------ */
/* kernel closures: electric.cpp:67-291 (a = source, b = target)            */

static double eval_poly(int kid, const double *a, const double *b) {
    /* polynomial surrogates with the same tensor structure as the PB
     * kernels (degree <= 2 per coordinate, no regularisation):
     * K1a = d_a, K2 = 0.5 + 0.25 r^2, K3ab = delta_ab + 0.5 d_a d_b,
     * K4a = -0.75 d_a, d = a - b. */
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    switch (kid) {
    case ORC_K1X: return d[0];
    case ORC_K1Y: return d[1];
    case ORC_K1Z: return d[2];
    case ORC_K2: return 0.5 + 0.25 * r2;
    case ORC_K3XX: return 1.0 + 0.5 * d[0] * d[0];
    case ORC_K3YY: return 1.0 + 0.5 * d[1] * d[1];
    case ORC_K3ZZ: return 1.0 + 0.5 * d[2] * d[2];
    case ORC_K3XY: case ORC_K3YX: return 0.5 * d[0] * d[1];
    case ORC_K3YZ: case ORC_K3ZY: return 0.5 * d[1] * d[2];
    case ORC_K3ZX: case ORC_K3XZ: return 0.5 * d[2] * d[0];
    case ORC_K4X: return -0.75 * d[0];
    case ORC_K4Y: return -0.75 * d[1];
    case ORC_K4Z: return -0.75 * d[2];
    }
    return 0.0;
}

double orc_kernel_eval(int kid, const orc_kparams *kp, const double *a, const double *b) {
    if (kp->mode == 1) return eval_poly(kid, a, b);
    const double kappa = kp->kappa, dE = kp->eps_out, dI = kp->eps_in, vr = kp->vacant_radius;
    double d = SQR(a[0] - b[0]) + SQR(a[1] - b[1]) + SQR(a[2] - b[2]);
    double r = sqrt(d);
    switch (kid) {
    case ORC_K1X: case ORC_K1Y: case ORC_K1Z: {
        int ax = kid - ORC_K1X;
        if (r < vr) return 0.;
        return (a[ax] - b[ax]) * (dE / dI * exp(-kappa * r) * (kappa * r + 1) - 1.) / 4.0 / ORC_PI / r / d;
    }
    case ORC_K2:
        if (r < vr) {
            if (kappa == 0.) return 0.;
            return (exp(-kappa * vr) - 1.0 + kappa * vr) / kappa / 2.0 / ORC_PI / vr / vr;
        }
        return (1.0 - exp(-kappa * r)) / 4.0 / ORC_PI / r;
    case ORC_K3XX: case ORC_K3YY: case ORC_K3ZZ: {
        int ax = kid - ORC_K3XX;
        double df = a[ax] - b[ax], t = kappa * r;
        if (r < vr) return 0.;
        return (1 - exp(-t) * (t + 1)) * 1.0 / d / r + (exp(-t) * ((t + 3) * t + 3) - 3.0) * df * df / d / d / r;
    }
    case ORC_K3XY: case ORC_K3YZ: case ORC_K3ZX: case ORC_K3YX: case ORC_K3ZY: case ORC_K3XZ: {
        /* (df1, df2) per closure: xy,yx -> (x,y); yz,zy -> (y,z); zx,xz -> (z,x) */
        int a1, a2;
        if (kid == ORC_K3XY || kid == ORC_K3YX) { a1 = 0; a2 = 1; }
        else if (kid == ORC_K3YZ || kid == ORC_K3ZY) { a1 = 1; a2 = 2; }
        else { a1 = 2; a2 = 0; }
        double t = kappa * r;
        if (r < vr) return 0.;
        return (exp(-t) * ((t + 3) * t + 3) - 3.0) * (a[a1] - b[a1]) * (a[a2] - b[a2]) / d / d / r;
    }
    case ORC_K4X: case ORC_K4Y: case ORC_K4Z: {
        int ax = kid - ORC_K4X;
        if (r < vr) return 0.;
        return -(a[ax] - b[ax]) * (dI / dE * exp(-kappa * r) * (kappa * r + 1) - 1.) / 4.0 / ORC_PI / r / d;
    }
    }
    return 0.0;
}

/* ------------------------------------------------------------------------ */
/* the FMM passes: bbfmm.h:500-623                                          */

typedef struct {
    const orc_tree *t;
    int p, r, kid;
    const orc_kparams *kp;
    double *cheb;   /* p standard nodes */
    double *R;      /* 8 r x r */
    double *M;      /* nnodes x r multipole (nodeCharge) */
    double *L;      /* nnodes x r local (nodePotential) */
} fmm_ctx;

/* scaledCnode tensor grid of a box (bbfmm.h:370-377, 456-467):
 * point k p^2 + j p + i = (cx + rx x_i, cy + ry x_j, cz + rz x_k) */
static void box_grid(const fmm_ctx *f, const orc_node *n, double *g) {
    int p = f->p;
    for (int k = 0; k < p; ++k)
        for (int j = 0; j < p; ++j)
            for (int i = 0; i < p; ++i) {
                double *q = g + 3 * (size_t)(k * p * p + j * p + i);
                q[0] = n->c[0] + n->r[0] * f->cheb[i];
                q[1] = n->c[1] + n->r[1] * f->cheb[j];
                q[2] = n->c[2] + n->r[2] * f->cheb[k];
            }
}

static void up_leaf(fmm_ctx *f, int id, const double *charge) {
    const orc_node *n = &f->t->nodes[id];
    double *M = f->M + (size_t)id * f->r;
    if (n->nsrc == 0) return; /* bbfmm.h:516-519 */
    double *S = (double *)malloc((size_t)n->nsrc * f->r * sizeof(double));
    leaf_matrix(f->p, f->t->src, n->src, n->nsrc, n->c, n->r, S);
    for (int m = 0; m < f->r; ++m) {
        double acc = 0.0;
        for (int k = 0; k < n->nsrc; ++k) acc += S[(size_t)m * n->nsrc + k] * charge[n->src[k]];
        M[m] += acc;
    }
    free(S);
}

static void up_internal(fmm_ctx *f, int id) {
    const orc_node *n = &f->t->nodes[id];
    double *M = f->M + (size_t)id * f->r;
    int r = f->r;
    for (int c = 0; c < 8; ++c) {
        const orc_node *cn = &f->t->nodes[n->child[c]];
        if (cn->empty) continue;
        const double *Mc = f->M + (size_t)n->child[c] * r;
        const double *Rc = f->R + (size_t)c * r * r;
        for (int col = 0; col < r; ++col) { /* M += R[c]^T Mc */
            double acc = 0.0;
            for (int row = 0; row < r; ++row) acc += Rc[(size_t)col * r + row] * Mc[row];
            M[col] += acc;
        }
    }
}

/* L_B += K(grid_A -> grid_B) M_A with K evaluated entry by entry
 * (bbfmm.h:547-561, kernelEvalChebyshev 456-479) */
static void m2l_pair(fmm_ctx *f, int a, int b, double *K, double *ga, double *gb) {
    const orc_tree *t = f->t;
    int r = f->r;
    box_grid(f, &t->nodes[a], ga);
    box_grid(f, &t->nodes[b], gb);
    for (int s = 0; s < r; ++s)
        for (int tt = 0; tt < r; ++tt) K[(size_t)s * r + tt] = orc_kernel_eval(f->kid, f->kp, ga + 3 * s, gb + 3 * tt);
    const double *Ma = f->M + (size_t)a * r;
    double *Lb = f->L + (size_t)b * r;
    for (int tt = 0; tt < r; ++tt) {
        double acc = 0.0;
        for (int s = 0; s < r; ++s) acc += K[(size_t)s * r + tt] * Ma[s];
        Lb[tt] += acc;
    }
}

static void down_m2l(fmm_ctx *f, int id) {
    const orc_tree *t = f->t;
    const orc_node *n = &t->nodes[id];
    if (n->parent == -1) return;
    int r = f->r;
    double *K = (double *)malloc((size_t)r * r * sizeof(double));
    double *ga = (double *)malloc(3 * (size_t)r * sizeof(double));
    double *gb = (double *)malloc(3 * (size_t)r * sizeof(double));
    for (int i = 0; i < n->nv; ++i)
        if (!t->nodes[n->v[i]].empty) m2l_pair(f, n->v[i], id, K, ga, gb);
    for (int i = 0; i < n->nx; ++i)
        if (!t->nodes[n->x[i]].empty) m2l_pair(f, n->x[i], id, K, ga, gb);
    free(K);
    free(ga);
    free(gb);
}

static void down_l2l(fmm_ctx *f, int id) {
    const orc_node *n = &f->t->nodes[id];
    if (n->parent == -1) return;
    int r = f->r;
    const double *Rc = f->R + (size_t)n->slot * r * r;
    const double *Lp = f->L + (size_t)n->parent * r;
    double *L = f->L + (size_t)id * r;
    for (int row = 0; row < r; ++row) { /* L += R[slot] Lp */
        double acc = 0.0;
        for (int col = 0; col < r; ++col) acc += Rc[(size_t)col * r + row] * Lp[col];
        L[row] += acc;
    }
}

/* U, W near field + L2P + scatter (bbfmm.h:571-610) */
static void down_leaf(fmm_ctx *f, int id, const double *charge, double *potential) {
    const orc_tree *t = f->t;
    const orc_node *n = &t->nodes[id];
    if (!n->leaf || n->ntgt == 0) return;
    int nt = n->ntgt, r = f->r;
    double *pot = (double *)calloc((size_t)nt, sizeof(double));
    for (int pass = 0; pass < 2; ++pass) {
        int len = pass == 0 ? n->nu : n->nw;
        const int *lst = pass == 0 ? n->u : n->w;
        for (int li = 0; li < len; ++li) {
            const orc_node *a = &t->nodes[lst[li]];
            if (a->empty) continue;
            /* K (nt x ns) then pot += K q_A */
            double *K = (double *)malloc((size_t)nt * a->nsrc * sizeof(double));
            for (int s = 0; s < a->nsrc; ++s)
                for (int k = 0; k < nt; ++k)
                    K[(size_t)s * nt + k] =
                        orc_kernel_eval(f->kid, f->kp, t->src + 3 * (size_t)a->src[s], t->tgt + 3 * (size_t)n->tgt[k]);
            for (int k = 0; k < nt; ++k) {
                double acc = 0.0;
                for (int s = 0; s < a->nsrc; ++s) acc += K[(size_t)s * nt + k] * charge[a->src[s]];
                pot[k] += acc;
            }
            free(K);
        }
    }
    double *Lm = (double *)malloc((size_t)nt * r * sizeof(double));
    leaf_matrix(f->p, t->tgt, n->tgt, nt, n->c, n->r, Lm);
    const double *L = f->L + (size_t)id * r;
    for (int k = 0; k < nt; ++k) {
        double acc = 0.0;
        for (int m = 0; m < r; ++m) acc += Lm[(size_t)m * nt + k] * L[m];
        pot[k] += acc;
    }
    for (int k = 0; k < nt; ++k) potential[n->tgt[k]] += pot[k];
    free(Lm);
    free(pot);
}

int orc_fmm_run(const orc_tree *t, int p, int kid, const orc_kparams *kp, const double *charge,
                double *potential) {
    fmm_ctx f;
    f.t = t;
    f.p = p;
    f.r = p * p * p;
    f.kid = kid;
    f.kp = kp;
    f.cheb = (double *)malloc((size_t)p * sizeof(double));
    orc_cheb_nodes(p, f.cheb);
    f.R = (double *)malloc(8 * (size_t)f.r * f.r * sizeof(double));
    orc_transfer_R(p, f.R);
    f.M = (double *)calloc((size_t)t->nnodes * f.r, sizeof(double));
    f.L = (double *)calloc((size_t)t->nnodes * f.r, sizeof(double));
    if (!f.cheb || !f.R || !f.M || !f.L) return -1;

    /* nodes grouped by level */
    int maxl = 0;
    for (int i = 0; i < t->nnodes; ++i)
        if (t->nodes[i].level > maxl) maxl = t->nodes[i].level;
    int *cnt = (int *)calloc((size_t)maxl + 2, sizeof(int));
    for (int i = 0; i < t->nnodes; ++i) cnt[t->nodes[i].level + 1]++;
    for (int l = 0; l <= maxl; ++l) cnt[l + 1] += cnt[l];
    int *bylvl = (int *)malloc((size_t)t->nnodes * sizeof(int));
    int *fill = (int *)calloc((size_t)maxl + 1, sizeof(int));
    for (int i = 0; i < t->nnodes; ++i) {
        int l = t->nodes[i].level;
        bylvl[cnt[l] + fill[l]++] = i;
    }

    /* upward pass, deepest level first */
    for (int l = maxl; l >= 0; --l) {
        int lo = cnt[l], hi = cnt[l + 1];
#pragma omp parallel for schedule(dynamic, 4)
        for (int q = lo; q < hi; ++q) {
            int id = bylvl[q];
            if (t->nodes[id].leaf) up_leaf(&f, id, charge);
            else up_internal(&f, id);
        }
    }
    /* downward pass: M2L (V, X) everywhere, then L2L top-down, then leaves */
    int nn = t->nnodes;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nn; ++i) down_m2l(&f, i);
    for (int l = 1; l <= maxl; ++l) {
        int lo = cnt[l], hi = cnt[l + 1];
#pragma omp parallel for schedule(static)
        for (int q = lo; q < hi; ++q) down_l2l(&f, bylvl[q]);
    }
    for (int i = 0; i < t->ntgt; ++i) potential[i] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nn; ++i) down_leaf(&f, i, charge, potential);

    free(cnt);
    free(bylvl);
    free(fill);
    free(f.cheb);
    free(f.R);
    free(f.M);
    free(f.L);
    return 0;
}

void orc_direct(int kid, const orc_kparams *kp, const double *src, int nsrc, const double *tgt, int ntgt,
                const double *charge, double *potential) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < ntgt; ++t) {
        double acc = 0.0;
        for (int s = 0; s < nsrc; ++s) acc += orc_kernel_eval(kid, kp, src + 3 * (size_t)s, tgt + 3 * (size_t)t) * charge[s];
        potential[t] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* matvec dmapping: electric.cpp:294-388                                    */

/* which weighted charge each kernel takes (electric.cpp:333-348):
 * 0 = phiX, 1 = phiY, 2 = phiZ, 3 = pphi_pn */
static const int kernel_charge[ORC_NKERNEL] = {0, 1, 2, 3, 0, 1, 2, 0, 1, 2, 1, 2, 0, 3, 3, 3};

static void form_charges(int n, const double *nrm, const double *w, const double *x, double *q) {
    for (int i = 0; i < n; ++i) {
        q[3 * (size_t)n + i] = x[n + i] * w[i];
        q[0 * (size_t)n + i] = x[i] * nrm[3 * (size_t)i + 0] * w[i];
        q[1 * (size_t)n + i] = x[i] * nrm[3 * (size_t)i + 1] * w[i];
        q[2 * (size_t)n + i] = x[i] * nrm[3 * (size_t)i + 2] * w[i];
    }
}

static void combine(int n, const double *nrm, const orc_kparams *kp, const double *x, double **ret, double *y) {
    const double dE = kp->eps_out, dI = kp->eps_in;
    for (int i = 0; i < n; ++i) {
        double nx = nrm[3 * (size_t)i], ny = nrm[3 * (size_t)i + 1], nz = nrm[3 * (size_t)i + 2];
        double tmp1 = nx * (ret[ORC_K3XX][i] + ret[ORC_K3YX][i] + ret[ORC_K3ZX][i]) +
                      ny * (ret[ORC_K3YY][i] + ret[ORC_K3XY][i] + ret[ORC_K3ZY][i]) +
                      nz * (ret[ORC_K3ZZ][i] + ret[ORC_K3XZ][i] + ret[ORC_K3YZ][i]);
        double tmp2 = nx * ret[ORC_K4X][i] + ny * ret[ORC_K4Y][i] + nz * ret[ORC_K4Z][i];
        y[i] = x[i] + (ret[ORC_K1X][i] + ret[ORC_K1Y][i] + ret[ORC_K1Z][i] - ret[ORC_K2][i]) / (0.5 * (1 + dE / dI));
        y[n + i] = x[n + i] + (tmp1 - tmp2) / (0.5 * (1 + dI / dE));
    }
}

int orc_dmapping(int n, const double *xyz, const double *nrm, const double *w, const orc_fmm_params *fp,
                 const double *x, double *y, int share_tree) {
    double *q = (double *)malloc(4 * (size_t)n * sizeof(double));
    double *buf = (double *)malloc(ORC_NKERNEL * (size_t)n * sizeof(double));
    double *ret[ORC_NKERNEL];
    form_charges(n, nrm, w, x, q);
    orc_tree shared;
    int rc = 0;
    if (share_tree && orc_tree_build(&shared, xyz, n, xyz, n, fp->max_point, fp->max_level) != 0) rc = -1;
    for (int k = 0; k < ORC_NKERNEL && rc == 0; ++k) {
        ret[k] = buf + (size_t)k * n;
        const double *ch = q + (size_t)kernel_charge[k] * n;
        if (share_tree) {
            rc = orc_fmm_run(&shared, fp->p, k, &fp->k, ch, ret[k]);
        } else {
            orc_tree t;
            if (orc_tree_build(&t, xyz, n, xyz, n, fp->max_point, fp->max_level) != 0) { rc = -1; break; }
            rc = orc_fmm_run(&t, fp->p, k, &fp->k, ch, ret[k]);
            orc_tree_free(&t);
        }
    }
    if (share_tree) orc_tree_free(&shared);
    if (rc == 0) combine(n, nrm, &fp->k, x, ret, y);
    free(q);
    free(buf);
    return rc;
}

void orc_dmapping_direct(int n, const double *xyz, const double *nrm, const double *w, const orc_kparams *kp,
                         const double *x, double *y) {
    double *q = (double *)malloc(4 * (size_t)n * sizeof(double));
    double *buf = (double *)malloc(ORC_NKERNEL * (size_t)n * sizeof(double));
    double *ret[ORC_NKERNEL];
    form_charges(n, nrm, w, x, q);
    for (int k = 0; k < ORC_NKERNEL; ++k) {
        ret[k] = buf + (size_t)k * n;
        orc_direct(k, kp, xyz, n, xyz, n, q + (size_t)kernel_charge[k] * n, ret[k]);
    }
    combine(n, nrm, kp, x, ret, y);
    free(q);
    free(buf);
}

/* RHS load (electric.cpp:395-413), rescale = 1 (coordinates in Angstrom) */
void orc_rhs(int n, const double *xyz, const double *nrm, int m, const double *atom_xyz, const double *atom_q,
             double eps_in, double eps_out, double *b) {
    const double dE = eps_out, dI = eps_in;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double b1 = 0.0, b2 = 0.0;
        const double *s = xyz + 3 * (size_t)i, *nv = nrm + 3 * (size_t)i;
        for (int a = 0; a < m; ++a) {
            const double *c = atom_xyz + 3 * (size_t)a;
            double d = SQR(s[0] - c[0]) + SQR(s[1] - c[1]) + SQR(s[2] - c[2]);
            double r = sqrt(d);
            b1 += atom_q[a] / dI / 4.0 / ORC_PI / r / (0.5 * (1 + dE / dI));
            b2 -= atom_q[a] / dI / 4.0 / ORC_PI / d / r *
                  (nv[0] * (s[0] - c[0]) + nv[1] * (s[1] - c[1]) + nv[2] * (s[2] - c[2])) / (0.5 * (1 + dI / dE));
        }
        b[i] = b1;
        b[n + i] = b2;
    }
}

/* polarizeMap (electric.cpp:442-485): output = K2[qpsi] - sum_a K1a[qa],
 * targets = atom centres, tree over the surface points */
int orc_polarize(int n, const double *xyz, const double *nrm, const double *w, int m, const double *atom_xyz,
                 const orc_fmm_params *fp, const double *x, double *phi_reac) {
    double *q = (double *)malloc(4 * (size_t)n * sizeof(double));
    double *ret = (double *)malloc(4 * (size_t)m * sizeof(double));
    form_charges(n, nrm, w, x, q);
    int kids[4] = {ORC_K1X, ORC_K1Y, ORC_K1Z, ORC_K2};
    int rc = 0;
    for (int k = 0; k < 4 && rc == 0; ++k) {
        orc_tree t;
        if (orc_tree_build(&t, xyz, n, atom_xyz, m, fp->max_point, fp->max_level) != 0) { rc = -1; break; }
        rc = orc_fmm_run(&t, fp->p, kids[k], &fp->k, q + (size_t)kernel_charge[kids[k]] * n, ret + (size_t)k * m);
        orc_tree_free(&t);
    }
    if (rc == 0)
        for (int i = 0; i < m; ++i) phi_reac[i] = ret[3 * (size_t)m + i] - (ret[i] + ret[(size_t)m + i] + ret[2 * (size_t)m + i]);
    free(q);
    free(ret);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* GMRES: gmres.cpp:10-164                                                  */

static double dot(int n, const double *a, const double *b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void gen_rot(double dx, double dy, double *cs, double *sn) { /* gmres.cpp:23-36 */
    if (dy == 0.0) {
        *cs = 1.0;
        *sn = 0.0;
    } else if (fabs(dy) > fabs(dx)) {
        double temp = dx / dy;
        *sn = 1.0 / sqrt(1.0 + temp * temp);
        *cs = temp * (*sn);
    } else {
        double temp = dy / dx;
        *cs = 1.0 / sqrt(1.0 + temp * temp);
        *sn = temp * (*cs);
    }
}

static void apply_rot(double *dx, double *dy, double cs, double sn) { /* gmres.cpp:39-44 */
    double temp = cs * (*dx) + sn * (*dy);
    *dy = -sn * (*dx) + cs * (*dy);
    *dx = temp;
}

/* gmres.cpp:10-21; H is (m+1) x (m+1) col-major */
static void update(int n, double *x, int k, const double *H, int ld, const double *s, double **v) {
    double *y = (double *)malloc((size_t)(k + 2) * sizeof(double));
    for (int i = 0; i <= k; ++i) y[i] = s[i];
    for (int i = k; i >= 0; i--) {
        y[i] /= H[(size_t)i * ld + i];
        for (int j = i - 1; j >= 0; j--) y[j] -= H[(size_t)i * ld + j] * y[i];
    }
    for (int j = 0; j <= k; j++)
        for (int q = 0; q < n; ++q) x[q] += y[j] * v[j][q];
    free(y);
}

int orc_gmres(orc_matvec_fn A, void *ctx, int n, const double *b, double *x, int m, int maxit, double tol,
              int *iters, double *resid, int *nmatvec) {
    int nmv = 0, inner = 0;
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double normb = sqrt(dot(n, b, b));
    A(ctx, x, p);
    nmv++;
    for (int q = 0; q < n; ++q) r[q] = b[q] - p[q];
    double beta = sqrt(dot(n, r, r));
    if (normb == 0.0) normb = 1;
    double res = beta / normb;
    int ret = 1;
    if (res <= tol) {
        ret = 0;
        goto done_early;
    }
    {
        int ld = m + 1;
        double **v = (double **)malloc((size_t)(m + 1) * sizeof(double *));
        for (int i = 0; i <= m; ++i) v[i] = (double *)malloc((size_t)n * sizeof(double));
        double *H = (double *)calloc((size_t)ld * ld, sizeof(double));
        double *s = (double *)calloc((size_t)m + 1, sizeof(double));
        double *cs = (double *)calloc((size_t)m + 1, sizeof(double));
        double *sn = (double *)calloc((size_t)m + 1, sizeof(double));
        double *w = (double *)malloc((size_t)n * sizeof(double));
        int i, j = 1;
        while (j <= maxit) {
            double ib = 1.0 / beta;
            for (int q = 0; q < n; ++q) v[0][q] = r[q] * ib;
            for (int q = 0; q <= m; ++q) s[q] = 0.0;
            s[0] = beta;
            for (i = 0; i < m && j <= maxit; i++, j++) {
                A(ctx, v[i], w);
                nmv++;
                inner++;
                for (int k = 0; k <= i; k++) {
                    double h = dot(n, w, v[k]);
                    H[(size_t)i * ld + k] = h;
                    for (int q = 0; q < n; ++q) w[q] -= h * v[k][q];
                }
                double hn = sqrt(dot(n, w, w));
                H[(size_t)i * ld + i + 1] = hn;
                double ih = 1.0 / hn;
                for (int q = 0; q < n; ++q) v[i + 1][q] = w[q] * ih;
                for (int k = 0; k < i; k++) apply_rot(&H[(size_t)i * ld + k], &H[(size_t)i * ld + k + 1], cs[k], sn[k]);
                gen_rot(H[(size_t)i * ld + i], H[(size_t)i * ld + i + 1], &cs[i], &sn[i]);
                apply_rot(&H[(size_t)i * ld + i], &H[(size_t)i * ld + i + 1], cs[i], sn[i]);
                apply_rot(&s[i], &s[i + 1], cs[i], sn[i]);
                if ((res = fabs(s[i + 1]) / normb) < tol) {
                    update(n, x, i, H, ld, s, v);
                    ret = 0;
                    goto done;
                }
            }
            update(n, x, i - 1, H, ld, s, v);
            A(ctx, x, p);
            nmv++;
            for (int q = 0; q < n; ++q) r[q] = b[q] - p[q];
            beta = sqrt(dot(n, r, r));
            if ((res = beta / normb) < tol) {
                ret = 0;
                goto done;
            }
        }
    done:
        for (int q = 0; q <= m; ++q) free(v[q]);
        free(v);
        free(H);
        free(s);
        free(cs);
        free(sn);
        free(w);
    }
done_early:
    free(r);
    free(p);
    if (iters) *iters = inner;
    if (resid) *resid = res;
    if (nmatvec) *nmatvec = nmv;
    return ret;
}

typedef struct {
    int n;
    const double *A;
} dense_ctx;

static void dense_mv(void *c, const double *x, double *y) {
    const dense_ctx *d = (const dense_ctx *)c;
    for (int i = 0; i < d->n; ++i) y[i] = 0.0;
    for (int j = 0; j < d->n; ++j)
        for (int i = 0; i < d->n; ++i) y[i] += d->A[(size_t)j * d->n + i] * x[j];
}

int orc_gmres_dense(int n, const double *Amat, const double *b, double *x, int m, int maxit, double tol, int *iters,
                    double *resid) {
    dense_ctx c = {n, Amat};
    return orc_gmres(dense_mv, &c, n, b, x, m, maxit, tol, iters, resid, NULL);
}

typedef struct {
    int n;
    const double *xyz, *nrm, *w;
    const orc_fmm_params *fp;
    int share_tree;
} fmm_op_ctx;

static void fmm_mv(void *c, const double *x, double *y) {
    const fmm_op_ctx *d = (const fmm_op_ctx *)c;
    orc_dmapping(d->n, d->xyz, d->nrm, d->w, d->fp, x, y, d->share_tree);
}

int orc_solve(int n, const double *xyz, const double *nrm, const double *w, const orc_fmm_params *fp,
              const double *b, double *x, int m, int maxit, double tol, int share_tree, int *iters, double *resid,
              int *nmatvec) {
    fmm_op_ctx c = {n, xyz, nrm, w, fp, share_tree};
    return orc_gmres(fmm_mv, &c, 2 * n, b, x, m, maxit, tol, iters, resid, nmatvec);
}
