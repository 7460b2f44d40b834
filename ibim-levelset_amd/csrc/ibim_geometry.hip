// ibim_geometry.hip -- the geometry pre-step on the GPU: the solvent-excluded
// surface as a level set on a uniform grid, and the IBIM tube quadrature
// nodes around it.
//
// Reference (levelset.cpp, main.cpp:47-62):
//   expand        max over atoms of (r + probe) - |x - c| in a window of
//                 (r + probe)/h + band cells (levelset.cpp:91-109)
//   evolve        TVD-RK3 of u_t = -v |grad u| with WENO5 one-sided
//                 derivatives and a Godunov norm, moving the surface in by
//                 the probe radius (levelset.cpp:111-187, 294-396)
//   reinitialize  TVD-RK3 of u_t = -sign(phi0)(|grad u| - 1) until every tube
//                 point has ||grad u| - 1| <= thres (levelset.cpp:191-287, 398-435)
//   Surface       exterior flood fill from grid point 0, cavities pushed out
//                 of the tube, then for each tube point the projection
//                 P - grad(phi) phi with the averaged one-sided gradient (not
//                 renormalised) and the cosine weight (levelset.cpp:437-586)
// Each grid-wide stage is one kernel; the grid is HBM resident throughout.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

#include "ibim_internal.h"

struct ibim_surface {
    int64_t n = 0;
    double *xyz = nullptr, *nrm = nullptr, *w = nullptr;
    ibim_surface_info info{};
};

namespace ibim {
namespace {

struct GridDims {
    int nx, ny, nz;
    double sx, sy, sz, h;
    double rh;  // 1 / h
};

inline int nblk(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

__device__ __forceinline__ double gget(const double *__restrict__ g, const GridDims &d, int i, int j, int k) {
    i = min(max(i, 0), d.nx - 1);  // clamped read (levelset.cpp:28-33)
    j = min(max(j, 0), d.ny - 1);
    k = min(max(k, 0), d.nz - 1);
    return g[((size_t)i * d.ny + j) * d.nz + k];
}

// WENO5 forward (up) and backward (un) derivatives (levelset.cpp:330-396)
// from the seven values w[l] at offsets l - 3 along one axis
__device__ __forceinline__ void weno7(const double (&w)[7], double rh, double &up, double &un) {
    // the divisions of the reference are multiplications by constant
    // reciprocals and Newton-refined reciprocals here (each within an ulp;
    // an FP64 division costs about four times as many instructions)
    const double EPS_LS = 1e-6;
    const double r12 = 1.0 / 12.0, r3 = 1.0 / 3.0, r6 = 1.0 / 6.0;
    double a = w[1] - w[0], b = w[2] - w[1], c = w[3] - w[2], dd = w[4] - w[3], e = w[5] - w[4], f = w[6] - w[5];
    const double df = (-b + 7.0 * (c + dd) - e) * r12;
    a = b - a;
    b = c - b;
    c = dd - c;
    dd = e - dd;
    e = f - e;
    double is0 = 13.0 * (e - dd) * (e - dd) + 3.0 * (e - 3.0 * dd) * (e - 3.0 * dd);
    double is1 = 13.0 * (dd - c) * (dd - c) + 3.0 * (dd + c) * (dd + c);
    double is2 = 13.0 * (c - b) * (c - b) + 3.0 * (3.0 * c - b) * (3.0 * c - b);
    // weights a_i / (a0 + a1 + a2) with a_i = (1, 6, 3)_i / q_i, q_i = (eps +
    // is_i)^2, formed as products of the q's over one reciprocal
    double q0 = (EPS_LS + is0) * (EPS_LS + is0), q1 = (EPS_LS + is1) * (EPS_LS + is1),
           q2 = (EPS_LS + is2) * (EPS_LS + is2);
    double p0 = q1 * q2, p2 = 3.0 * q0 * q1;
    double rs = recip(p0 + 6.0 * q0 * q2 + p2);
    double w0 = p0 * rs, w2 = p2 * rs;
    up = (df + w0 * (e - 2 * dd + c) * r3 + (w2 - 0.5) * (dd - 2.0 * c + b) * r6) * rh;
    is0 = 13.0 * (a - b) * (a - b) + 3.0 * (a - 3.0 * b) * (a - 3.0 * b);
    is1 = 13.0 * (b - c) * (b - c) + 3.0 * (b + c) * (b + c);
    is2 = 13.0 * (c - dd) * (c - dd) + 3.0 * (3.0 * c - dd) * (3.0 * c - dd);
    q0 = (EPS_LS + is0) * (EPS_LS + is0);
    q1 = (EPS_LS + is1) * (EPS_LS + is1);
    q2 = (EPS_LS + is2) * (EPS_LS + is2);
    p0 = q1 * q2;
    p2 = 3.0 * q0 * q1;
    rs = recip(p0 + 6.0 * q0 * q2 + p2);
    w0 = p0 * rs;
    w2 = p2 * rs;
    un = (df - w0 * (a - 2 * b + c) * r3 - (w2 - 0.5) * (b - 2 * c + dd) * r6) * rh;
}

// the same along dir at grid point (i, j, k), clamped reads
__device__ __forceinline__ void weno(const double *__restrict__ g, const GridDims &d, int i, int j, int k, int dir,
                                     double &up, double &un) {
    double w[7];
#pragma unroll
    for (int l = 0; l < 7; ++l) {
        const int o = l - 3;
        w[l] = gget(g, d, i + (dir == 0 ? o : 0), j + (dir == 1 ? o : 0), k + (dir == 2 ? o : 0));
    }
    weno7(w, d.rh, up, un);
}

// Godunov norm getNorm(A, B) = sqrt(sum max(A+, B-)^2) (levelset.cpp:294-301)
__device__ __forceinline__ double godunov(const double *A, const double *B) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double t = fmax(A[q] > 0. ? A[q] : 0., B[q] < 0. ? -B[q] : 0.);
        v += t * t;
    }
    return sqrt(v);
}

__device__ __forceinline__ void index3(int64_t I, const GridDims &d, int &i, int &j, int &k) {
    k = (int)(I % d.nz);
    const int64_t t = I / d.nz;
    j = (int)(t % d.ny);
    i = (int)(t / d.ny);
}

__global__ void k_fill(double *__restrict__ g, int64_t n, double v) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) g[I] = v;
}

// order-preserving map of doubles onto unsigned integers for atomicMax
__device__ __forceinline__ unsigned long long ord(double v) {
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unord(unsigned long long u) {
    u = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    return __longlong_as_double((long long)u);
}

__global__ void k_to_ord(double *__restrict__ g, int64_t n) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) {
        unsigned long long u = ord(g[I]);
        g[I] = __longlong_as_double((long long)u);
    }
}
__global__ void k_from_ord(double *__restrict__ g, int64_t n) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) g[I] = unord((unsigned long long)__double_as_longlong(g[I]));
}

// expand: one block row per atom, window cells over blockIdx.x (levelset.cpp:91-109)
__global__ void k_expand(unsigned long long *__restrict__ g, GridDims d, const double *__restrict__ axyz,
                         const double *__restrict__ rad, int m, double probe, int band, int wmax) {
    const int a = blockIdx.y + gridDim.y * blockIdx.z;
    if (a >= m) return;
    const double cx = axyz[3 * (size_t)a], cy = axyz[3 * (size_t)a + 1], cz = axyz[3 * (size_t)a + 2];
    const double rp = rad[a] + probe;
    const int r = (int)(rp / d.h) + band;
    const int icx = (int)((cx - d.sx) / d.h), icy = (int)((cy - d.sy) / d.h), icz = (int)((cz - d.sz) / d.h);
    const int w = 2 * r + 1;
    const int64_t cells = (int64_t)w * w * w;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cells; t += (int64_t)gridDim.x * blockDim.x) {
        const int oc = (int)(t % w), ob = (int)((t / w) % w), oa = (int)(t / ((int64_t)w * w));
        const int ia = icx - r + oa, ib = icy - r + ob, ic = icz - r + oc;
        // the reference writes without a bounds check; cells outside the grid are skipped here
        if (ia < 0 || ib < 0 || ic < 0 || ia >= d.nx || ib >= d.ny || ic >= d.nz) continue;
        const double px = d.sx + ia * d.h - cx, py = d.sy + ib * d.h - cy, pz = d.sz + ic * d.h - cz;
        const double dist = rp - sqrt(px * px + py * py + pz * pz);
        atomicMax(&g[((size_t)ia * d.ny + ib) * d.nz + ic], ord(dist));
    }
    (void)wmax;
}

// one TVD-RK3 stage of evolve (MODE 0) or reinitialize (MODE 1)
//   stage 1: out = base - T(base)
//   stage 2: out = (3 base + src - T(src)) / 4
//   stage 3: out = (base + 2 (src - T(src))) / 3      (out may alias base)
// T(u) = dtv |grad u|_G (evolve) or dtv sgn (|grad u|_G - 1) (reinit, sgn =
// phi0 / sqrt(phi0^2 + h^2) per point, k_sign).  A block is a tile of 64 (k)
// x 4 (j) points of one i-plane: the source's rows j0 - 3 .. j0 + 6 and
// columns k0 - 3 .. k0 + 66 of the plane are staged in LDS, clamped at the
// grid's faces as gget clamps, and the j and k stencils read them there; the
// i stencil reads the seven planes directly through plane pointers that are
// uniform over the block, at a 32-bit in-plane byte offset.  Needs
// ny * nz * 8 < 2^32 (the host checks).
constexpr int kLsTk = 64, kLsTj = 4;
constexpr int kLsRows = kLsTj + 6, kLsCols = kLsTk + 6;

__device__ __forceinline__ double at_byte(const double *p, unsigned off) {
    return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(p) + off);
}

// one point of k_ls_stage (its tile's rows and columns in t), and with
// COUNT the point's contribution to the gradient count
template <int MODE, int STAGE, bool COUNT>
__device__ __forceinline__ void ls_point(const double *base, const double *__restrict__ src,
                                         const double *__restrict__ sgn, double *out, const GridDims &d, double dtv,
                                         const double (&t)[kLsRows][kLsCols], int i, int j, int k, int tj, int tk,
                                         double tube, double thres, unsigned int &c0, unsigned int &c1, double &e) {
    const size_t plane = (size_t)d.ny * d.nz;
    const unsigned off = 8u * (unsigned)(j * d.nz + k);
    double up[3], un[3];
    {
        double w[7];
#pragma unroll
        for (int l = 0; l < 7; ++l) {
            const int ii = min(max(i + l - 3, 0), d.nx - 1);
            w[l] = l == 3 ? t[tj + 3][tk + 3] : at_byte(src + (size_t)ii * plane, off);
        }
        weno7(w, d.rh, up[0], un[0]);
    }
    {
        double w[7];
#pragma unroll
        for (int l = 0; l < 7; ++l) w[l] = t[tj + l][tk + 3];
        weno7(w, d.rh, up[1], un[1]);
    }
    {
        double w[7];
#pragma unroll
        for (int l = 0; l < 7; ++l) w[l] = t[tj + 3][tk + l];
        weno7(w, d.rh, up[2], un[2]);
    }
    const double c = t[tj + 3][tk + 3];  // src at the point
    double T;
    if constexpr (MODE == 0) {
        T = dtv * godunov(un, up);
    } else {
        const double s = at_byte(sgn + (size_t)i * plane, off);
        const double nd = s > 0. ? godunov(un, up) : godunov(up, un);
        T = dtv * s * (nd - 1.0);
        if constexpr (COUNT) {
            if (fabs(c) < tube) {  // a tube point: |grad| by getNorm, bad beyond thres
                const double nr = s > 0. ? nd : godunov(un, up);
                c0 = 1;
                c1 = fabs(nr - 1.0) > thres ? 1 : 0;
                e = fabs(nr - 1.0);
            }
        }
    }
    double *op = out + (size_t)i * plane;
    double v;
    if constexpr (STAGE == 1) {
        v = c - T;  // base is src
    } else {
        const double b = at_byte(base + (size_t)i * plane, off);
        if constexpr (STAGE == 2) v = (3 * b + c - T) * 0.25;
        else v = (b + 2 * (c - T)) * (1.0 / 3.0);
    }
    *reinterpret_cast<double *>(reinterpret_cast<char *>(op) + off) = v;
}

// COUNT (the first stage of a reinitialisation step): also countGradient
// (levelset.cpp:398-435) of src, the state the step starts from: tube points,
// bad points and the sum of ||grad| - 1| over the tube, reduced by wave
// shuffles and LDS to one slot per tile in tile order
template <int MODE, int STAGE, bool COUNT = false>
__global__ __launch_bounds__(256) void k_ls_stage(const double *base, const double *__restrict__ src,
                                                  const double *__restrict__ sgn, double *out, GridDims d,
                                                  double dtv, double tube = 0.0, double thres = 0.0,
                                                  unsigned int *__restrict__ pc = nullptr,
                                                  double *__restrict__ pe = nullptr) {
    __shared__ double t[kLsRows][kLsCols];
    const int k0 = blockIdx.x * kLsTk, j0 = blockIdx.y * kLsTj, i = blockIdx.z;
    const size_t plane = (size_t)d.ny * d.nz;
    const double *sp = src + (size_t)i * plane;
    for (int e = threadIdx.y * kLsTk + threadIdx.x; e < kLsRows * kLsCols; e += kLsTk * kLsTj) {
        const int r = e / kLsCols, c = e - r * kLsCols;
        const int jj = min(max(j0 - 3 + r, 0), d.ny - 1);
        const int kk = min(max(k0 - 3 + c, 0), d.nz - 1);
        t[r][c] = at_byte(sp, 8u * (unsigned)(jj * d.nz + kk));
    }
    __syncthreads();
    const int tk = threadIdx.x, tj = threadIdx.y;
    const int k = k0 + tk, j = j0 + tj;
    if constexpr (COUNT) {
        // every thread reaches the block's count below
        unsigned int c0 = 0, c1 = 0;
        double e = 0.0;
        if (k < d.nz && j < d.ny) ls_point<MODE, STAGE, true>(base, src, sgn, out, d, dtv, t, i, j, k, tj, tk, tube,
                                                                thres, c0, c1, e);
        for (int o = 32; o; o >>= 1) {
            c0 += __shfl_xor(c0, o);
            c1 += __shfl_xor(c1, o);
            e += __shfl_xor(e, o);
        }
        __shared__ unsigned int s0[4], s1[4];
        __shared__ double se[4];
        const int w = (tj * kLsTk + tk) >> 6;
        if (((tj * kLsTk + tk) & 63) == 0) {
            s0[w] = c0;
            s1[w] = c1;
            se[w] = e;
        }
        __syncthreads();
        if (tj == 0 && tk == 0) {
            const size_t b = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
            pc[2 * b] = s0[0] + s0[1] + s0[2] + s0[3];
            pc[2 * b + 1] = s1[0] + s1[1] + s1[2] + s1[3];
            pe[b] = (se[0] + se[1]) + (se[2] + se[3]);
        }
        return;
    } else {
        if (k >= d.nz || j >= d.ny) return;
        unsigned int c0, c1;
        double e;
        ls_point<MODE, STAGE, false>(base, src, sgn, out, d, dtv, t, i, j, k, tj, tk, 0.0, 0.0, c0, c1, e);
    }
}


// phi0 -> sign(phi0) = phi0 / sqrt(phi0^2 + eps) in place (the reinitialisation's
// smoothed sign, levelset.cpp:191-287; constant over its steps)
__global__ void k_sign(double *__restrict__ g, int64_t n, double eps) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) {
        const double p0 = g[I];
        g[I] = p0 / sqrt(p0 * p0 + eps);
    }
}

__global__ void k_scale(double *__restrict__ g, int64_t n, double s) {
    int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) g[I] *= s;
}

// one block of 1024: the per-tile slots of the counting stage summed in tile
// order (exact counts, a deterministic error sum)
__global__ __launch_bounds__(1024) void k_count_sum(int nb, const unsigned int *__restrict__ pc,
                                                    const double *__restrict__ pe,
                                                    unsigned long long *__restrict__ counts, double *__restrict__ err) {
    unsigned long long c0 = 0, c1 = 0;
    double e = 0.0;
    for (int b = threadIdx.x; b < nb; b += 1024) {
        c0 += pc[2 * (size_t)b];
        c1 += pc[2 * (size_t)b + 1];
        e += pe[b];
    }
    for (int o = 32; o; o >>= 1) {
        c0 += __shfl_xor(c0, o);
        c1 += __shfl_xor(c1, o);
        e += __shfl_xor(e, o);
    }
    __shared__ unsigned long long s0[16], s1[16];
    __shared__ double se[16];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s0[w] = c0;
        s1[w] = c1;
        se[w] = e;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t0 = 0, t1 = 0;
        double te = 0.0;
        for (int q = 0; q < 16; ++q) {
            t0 += s0[q];
            t1 += s1[q];
            te += se[q];
        }
        counts[0] = t0;
        counts[1] = t1;
        *err = te;
    }
}

// exterior flood fill (levelset.cpp:444-537): a point is reached when a
// reached neighbour has phi > -tube.  Computed on 8 x 8 x 8 tiles: a block
// loads its tile and a one-cell halo (reached flags and phi > -tube) into LDS
// and propagates inside the tile until nothing changes, so one launch
// carries the front across a whole tile; the host repeats launches until no
// tile changes.  The reached set is the fixpoint of the rule, whatever the
// order of updates.
__global__ __launch_bounds__(512) void k_flood_tile(const double *__restrict__ g, GridDims d, unsigned char *vis,
                                                    double neg_tube, int *__restrict__ changed) {
    constexpr int T = 8, H = T + 2;
    __shared__ unsigned char sv[H * H * H], sp[H * H * H];
    __shared__ int chg;
    const int ntx = (d.nx + T - 1) / T, nty = (d.ny + T - 1) / T, ntz = (d.nz + T - 1) / T;
    const int b = blockIdx.x;
    const int tk = b % ntz, tj = (b / ntz) % nty, ti = b / (ntz * nty);
    if (ti >= ntx) return;
    const int i0 = ti * T - 1, j0 = tj * T - 1, k0 = tk * T - 1;
    for (int c = threadIdx.x; c < H * H * H; c += blockDim.x) {
        const int ck = c % H, cj = (c / H) % H, ci = c / (H * H);
        const int i = i0 + ci, j = j0 + cj, k = k0 + ck;
        unsigned char v = 0, pass = 0;
        if (i >= 0 && i < d.nx && j >= 0 && j < d.ny && k >= 0 && k < d.nz) {
            const int64_t I = ((int64_t)i * d.ny + j) * d.nz + k;
            v = vis[I];
            pass = g[I] > neg_tube ? 1 : 0;
        }
        sv[c] = v;
        sp[c] = pass;
    }
    // this thread's interior cell
    const int kk = threadIdx.x % T, jj = (threadIdx.x / T) % T, ii = threadIdx.x / (T * T);
    const int i = ti * T + ii, j = tj * T + jj, k = tk * T + kk;
    const bool inside = i < d.nx && j < d.ny && k < d.nz;
    const int c = ((ii + 1) * H + (jj + 1)) * H + (kk + 1);
    bool mine = false;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) chg = 0;
        __syncthreads();
        if (inside && !sv[c] &&
            ((sv[c - H * H] && sp[c - H * H]) || (sv[c + H * H] && sp[c + H * H]) || (sv[c - H] && sp[c - H]) ||
             (sv[c + H] && sp[c + H]) || (sv[c - 1] && sp[c - 1]) || (sv[c + 1] && sp[c + 1]))) {
            sv[c] = 1;
            mine = true;
            chg = 1;
        }
        __syncthreads();
        if (!chg) break;
    }
    if (mine) {
        vis[((int64_t)i * d.ny + j) * d.nz + k] = 1;
        *changed = 1;
    }
}

__global__ void k_inclusion(double *__restrict__ g, int64_t n, const unsigned char *__restrict__ vis, double tube,
                            unsigned long long *__restrict__ cnt) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n) return;
    if (!vis[I] && fabs(g[I]) < tube) {
        g[I] = -(tube + 1e-15);
        atomicAdd(cnt, 1ull);
    }
}

__global__ void k_tube_flag(const double *__restrict__ g, int64_t n, double tube, int *__restrict__ flag) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I < n) flag[I] = fabs(g[I]) < tube ? 1 : 0;
}

// tube nodes, normals and weights in grid order (levelset.cpp:540-583),
// converted as electric.cpp:17-41 (nodes / s, weights * s * (h/s)^3)
__global__ __launch_bounds__(256) void k_tube_emit(const double *__restrict__ g, GridDims d,
                                                   const int *__restrict__ flag, const int *__restrict__ pos,
                                                   double tube, double inv_s, double wscale, double *__restrict__ xyz,
                                                   double *__restrict__ nrm, double *__restrict__ w) {
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= n || !flag[I]) return;
    int i, j, k;
    index3(I, d, i, j, k);
    double up[3], un[3], gr[3];
#pragma unroll
    for (int dir = 0; dir < 3; ++dir) {
        weno(g, d, i, j, k, dir, up[dir], un[dir]);
        gr[dir] = (un[dir] + up[dir]) * 0.5;
    }
    const double dist = gget(g, d, i, j, k);
    const double P[3] = {d.sx + i * d.h, d.sy + j * d.h, d.sz + k * d.h};
    const int o = pos[I];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        xyz[3 * (size_t)o + q] = (P[q] - gr[q] * dist) * inv_s;
        nrm[3 * (size_t)o + q] = gr[q];
    }
    w[o] = 0.5 * (1.0 + cos(M_PI * dist / tube)) / tube * wscale;
}

}  // namespace

// --------------------------------------------------------------------------
struct LsRun {
    GridDims d{};
    int64_t n = 0;
    double *g = nullptr, *phi0 = nullptr, *u1 = nullptr, *u2 = nullptr;
    hipStream_t st = nullptr;
};

static int ls_rk3(LsRun &L, int mode, double final_t, double vel, double cfl, double tube, double thres,
                  int *steps_out, double *quality) {
    double dt = cfl * L.d.h / vel;
    const int num_steps = (int)(final_t / dt) + 1;
    dt = final_t / (double)num_steps;
    const double dtv = dt * vel;
    const double eps = L.d.h * L.d.h;
    const int nb = nblk(L.n);
    // the tiles of k_ls_stage: 64 (k) x 4 (j) points of one i-plane
    const dim3 tgrid((L.d.nz + kLsTk - 1) / kLsTk, (L.d.ny + kLsTj - 1) / kLsTj, L.d.nx), tblk(kLsTk, kLsTj);
    const size_t nbt = (size_t)tgrid.x * tgrid.y * tgrid.z;  // count slots: one per tile
    if (nbt > (size_t)INT_MAX) {
        set_error("level set: more tiles than the gradient count's slots can index");
        return IBIM_ERR_ARG;
    }
    // the count's slots, its sums (two counts and the error sum, one
    // device block) and their pinned host copy, released on every return
    struct CountBufs {
        unsigned int *pc = nullptr;
        double *pe = nullptr;
        unsigned long long *res = nullptr;
        unsigned long long *host = nullptr;
        ~CountBufs() {
            for (void *p : {(void *)pc, (void *)pe, (void *)res})
                if (p) (void)hipFree(p);
            if (host) (void)hipHostFree(host);
        }
    } cb;
    IBIM_HIP_CHECK(dev_malloc(&cb.pc, 2 * nbt * sizeof(unsigned int)));
    IBIM_HIP_CHECK(dev_malloc(&cb.pe, nbt * sizeof(double)));
    IBIM_HIP_CHECK(dev_malloc(&cb.res, 3 * sizeof(unsigned long long)));
    IBIM_HIP_CHECK(hipHostMalloc((void **)&cb.host, 3 * sizeof(unsigned long long)));
    // the slots the first stage of a step wrote for the state it started
    // from, summed in tile order and read back in one copy
    auto count = [&](unsigned long long *h) -> int {
        k_count_sum<<<1, 1024, 0, L.st>>>((int)nbt, cb.pc, cb.pe, cb.res, (double *)(cb.res + 2));
        IBIM_LAUNCH_CHECK();
        IBIM_HIP_CHECK(hipMemcpyAsync(cb.host, cb.res, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, L.st));
        IBIM_HIP_CHECK(hipStreamSynchronize(L.st));
        h[0] = cb.host[0];
        h[1] = cb.host[1];
        double e;
        std::memcpy(&e, cb.host + 2, sizeof(double));
        if (quality) {
            quality[0] = h[0] ? e / (double)h[0] : 0.0;
            quality[1] = h[0] ? (double)h[1] / (double)h[0] : 0.0;
        }
        return IBIM_OK;
    };
    int step = 0, rc = IBIM_OK;
    unsigned long long hc[2] = {0, 1};
    if (mode == 0) {
        while (step < num_steps) {
            ++step;
            k_ls_stage<0, 1><<<tgrid, tblk, 0, L.st>>>(L.g, L.g, nullptr, L.u1, L.d, dtv);
            k_ls_stage<0, 2><<<tgrid, tblk, 0, L.st>>>(L.g, L.u1, nullptr, L.u2, L.d, dtv);
            k_ls_stage<0, 3><<<tgrid, tblk, 0, L.st>>>(L.g, L.u2, nullptr, L.g, L.d, dtv);
            IBIM_LAUNCH_CHECK();
        }
    } else {
        k_sign<<<nb, 256, 0, L.st>>>(L.phi0, L.n, eps);
        // Steps run while the state has a bad tube point and steps are left
        // (the gradient count before every step and after the last one).
        // Each step's first stage counts the state it starts from, so the
        // count is no pass of its own; the first stage after the last step
        // only counts (its result is not used).
        for (;;) {
            k_ls_stage<1, 1, true><<<tgrid, tblk, 0, L.st>>>(L.g, L.g, L.phi0, L.u1, L.d, dtv, tube, thres, cb.pc,
                                                             cb.pe);
            IBIM_LAUNCH_CHECK();
            if ((rc = count(hc))) return rc;
            if (step >= num_steps || hc[1] == 0) break;
            ++step;
            k_ls_stage<1, 2><<<tgrid, tblk, 0, L.st>>>(L.g, L.u1, L.phi0, L.u2, L.d, dtv);
            k_ls_stage<1, 3><<<tgrid, tblk, 0, L.st>>>(L.g, L.u2, L.phi0, L.g, L.d, dtv);
            IBIM_LAUNCH_CHECK();
        }
    }
    if (mode == 0) IBIM_HIP_CHECK(hipStreamSynchronize(L.st));
    *steps_out = step;
    return IBIM_OK;
}

// Surface: flood fill, inclusions, tube extraction into a new ibim_surface
static int ls_surface(LsRun &L, const ibim_geom_params &gp, ibim_surface *S) {
    const double tube = gp.thickness * L.d.h;
    const int nb = nblk(L.n);
    unsigned char *vis = nullptr;
    int *changed = nullptr;
    IBIM_HIP_CHECK(dev_malloc(&vis, (size_t)L.n));
    IBIM_HIP_CHECK(dev_malloc(&changed, sizeof(int)));
    IBIM_HIP_CHECK(hipMemsetAsync(vis, 0, (size_t)L.n, L.st));
    const unsigned char one = 1;
    IBIM_HIP_CHECK(hipMemcpyAsync(vis, &one, 1, hipMemcpyHostToDevice, L.st));
    const int ntiles = ((L.d.nx + 7) / 8) * ((L.d.ny + 7) / 8) * ((L.d.nz + 7) / 8);
    for (int it = 0;; ++it) {
        IBIM_HIP_CHECK(hipMemsetAsync(changed, 0, sizeof(int), L.st));
        for (int r = 0; r < 2; ++r) k_flood_tile<<<ntiles, 512, 0, L.st>>>(L.g, L.d, vis, -tube, changed);
        IBIM_LAUNCH_CHECK();
        int hch = 0;
        IBIM_HIP_CHECK(hipMemcpyAsync(&hch, changed, sizeof(int), hipMemcpyDeviceToHost, L.st));
        IBIM_HIP_CHECK(hipStreamSynchronize(L.st));
        if (!hch) break;
        if (it > 1000000) {
            set_error("flood fill did not terminate");
            return IBIM_ERR_STATE;
        }
    }
    unsigned long long *cnt = nullptr;
    IBIM_HIP_CHECK(dev_malloc(&cnt, sizeof(unsigned long long)));
    IBIM_HIP_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), L.st));
    k_inclusion<<<nb, 256, 0, L.st>>>(L.g, L.n, vis, tube, cnt);
    IBIM_LAUNCH_CHECK();
    unsigned long long hinc = 0;
    IBIM_HIP_CHECK(hipMemcpyAsync(&hinc, cnt, sizeof(hinc), hipMemcpyDeviceToHost, L.st));
    (void)hipFree(vis);
    (void)hipFree(changed);
    (void)hipFree(cnt);

    int *flag = nullptr, *pos = nullptr;
    IBIM_HIP_CHECK(dev_malloc(&flag, ((size_t)L.n + 1) * sizeof(int)));
    IBIM_HIP_CHECK(dev_malloc(&pos, ((size_t)L.n + 1) * sizeof(int)));
    IBIM_HIP_CHECK(hipMemsetAsync(flag, 0, ((size_t)L.n + 1) * sizeof(int), L.st));
    k_tube_flag<<<nb, 256, 0, L.st>>>(L.g, L.n, tube, flag);
    IBIM_LAUNCH_CHECK();
    size_t bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flag, pos, (int)(L.n + 1), L.st));
    void *tmp = nullptr;
    IBIM_HIP_CHECK(dev_malloc(&tmp, bytes ? bytes : 1));
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, flag, pos, (int)(L.n + 1), L.st));
    int total = 0;
    IBIM_HIP_CHECK(hipMemcpyAsync(&total, pos + L.n, sizeof(int), hipMemcpyDeviceToHost, L.st));
    IBIM_HIP_CHECK(hipStreamSynchronize(L.st));
    (void)hipFree(tmp);
    if (total <= 0) {
        (void)hipFree(flag);
        (void)hipFree(pos);
        set_error("the level set has no tube points");
        return IBIM_ERR_STATE;
    }
    S->n = total;
    IBIM_HIP_CHECK(dev_malloc(&S->xyz, 3 * (size_t)total * sizeof(double)));
    IBIM_HIP_CHECK(dev_malloc(&S->nrm, 3 * (size_t)total * sizeof(double)));
    IBIM_HIP_CHECK(dev_malloc(&S->w, (size_t)total * sizeof(double)));
    const double s = gp.scale > 0 ? gp.scale : 1.0;
    const double hs = L.d.h / s;
    k_tube_emit<<<nb, 256, 0, L.st>>>(L.g, L.d, flag, pos, tube, 1.0 / s, s * hs * hs * hs, S->xyz, S->nrm, S->w);
    IBIM_LAUNCH_CHECK();
    IBIM_HIP_CHECK(hipStreamSynchronize(L.st));
    (void)hipFree(flag);
    (void)hipFree(pos);
    std::vector<double> hw(total);
    IBIM_HIP_CHECK(hipMemcpy(hw.data(), S->w, (size_t)total * sizeof(double), hipMemcpyDeviceToHost));
    double area = 0.0;
    for (double v : hw) area += v;
    ibim_surface_info &inf = S->info;
    inf.n_nodes = total;
    inf.dims[0] = L.d.nx;
    inf.dims[1] = L.d.ny;
    inf.dims[2] = L.d.nz;
    inf.lo[0] = L.d.sx;
    inf.lo[1] = L.d.sy;
    inf.lo[2] = L.d.sz;
    inf.h = L.d.h;
    inf.inclusions = (int64_t)hinc;
    inf.area = area;
    return IBIM_OK;
}

static int check_geom(const ibim_geom_params *gp) {
    if (!gp || !(gp->h > 0) || gp->thickness < 1 || gp->band < 0 || !(gp->probe >= 0) || !(gp->reinit_time >= 0) ||
        !(gp->scale >= 0) || !(gp->init <= 0)) {
        set_error("bad geometry parameters (h > 0, thickness >= 1, band >= 0, probe >= 0, init <= 0 required)");
        return IBIM_ERR_ARG;
    }
    return IBIM_OK;
}

static int alloc_grid(LsRun &L, bool full) {
    L.n = (int64_t)L.d.nx * L.d.ny * L.d.nz;
    if (L.d.nx < 7 || L.d.ny < 7 || L.d.nz < 7 || L.n > ((int64_t)1 << 31) - 2 ||
        (int64_t)L.d.ny * L.d.nz > ((int64_t)1 << 29) - 1 || L.d.nx > 65535) {
        set_error("grid dimensions must be >= 7 and hold < 2^31 points, < 2^29 in a plane, <= 65535 planes");
        return IBIM_ERR_ARG;
    }
    IBIM_HIP_CHECK(dev_malloc(&L.g, (size_t)L.n * sizeof(double)));
    if (full) {
        IBIM_HIP_CHECK(dev_malloc(&L.phi0, (size_t)L.n * sizeof(double)));
        IBIM_HIP_CHECK(dev_malloc(&L.u1, (size_t)L.n * sizeof(double)));
        IBIM_HIP_CHECK(dev_malloc(&L.u2, (size_t)L.n * sizeof(double)));
    }
    return IBIM_OK;
}

static void free_grid(LsRun &L) {
    for (double *p : {L.g, L.phi0, L.u1, L.u2})
        if (p) (void)hipFree(p);
    if (L.st) (void)hipStreamDestroy(L.st);
}

}  // namespace ibim

using namespace ibim;

extern "C" {

void ibim_surface_destroy(ibim_surface *s) {
    if (!s) return;
    for (double *p : {s->xyz, s->nrm, s->w})
        if (p) (void)hipFree(p);
    delete s;
}

int ibim_surface_from_atoms(const ibim_geom_params *gp, int64_t m, const double *axyz, const double *radii,
                            ibim_surface **out) {
    if (!out || m < 1 || !axyz || !radii) {
        set_error("bad argument to ibim_surface_from_atoms");
        return IBIM_ERR_ARG;
    }
    *out = nullptr;
    int rc = check_geom(gp);
    if (rc) return rc;
    LsRun L;
    L.d.h = gp->h;
    L.d.rh = 1.0 / gp->h;
    if (gp->dims[0] > 0 && gp->dims[1] > 0 && gp->dims[2] > 0) {
        L.d.nx = gp->dims[0];
        L.d.ny = gp->dims[1];
        L.d.nz = gp->dims[2];
        L.d.sx = gp->lo[0];
        L.d.sy = gp->lo[1];
        L.d.sz = gp->lo[2];
    } else {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t a = 0; a < m; ++a)
            for (int q = 0; q < 3; ++q) {
                lo[q] = std::min(lo[q], axyz[3 * a + q] - radii[a] - gp->probe);
                hi[q] = std::max(hi[q], axyz[3 * a + q] + radii[a] + gp->probe);
            }
        int *dd[3] = {&L.d.nx, &L.d.ny, &L.d.nz};
        double *ss[3] = {&L.d.sx, &L.d.sy, &L.d.sz};
        for (int q = 0; q < 3; ++q) {
            *ss[q] = lo[q] - gp->pad;
            *dd[q] = (int)std::ceil((hi[q] + gp->pad - *ss[q]) / gp->h) + 1;
        }
    }
    IBIM_HIP_CHECK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
    if ((rc = alloc_grid(L, true))) {
        free_grid(L);
        return rc;
    }
    ibim_surface *S = new ibim_surface();
    const int nb = nblk(L.n);
    double *d_at = nullptr;
    auto fail = [&](int code) {
        if (d_at) (void)hipFree(d_at);
        free_grid(L);
        ibim_surface_destroy(S);
        return code;
    };
    // expand (levelset.cpp:91-109) onto the initial value (main.cpp:43: -2 * size)
    if (dev_malloc(&d_at, 4 * (size_t)m * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc (atoms)");
        return fail(IBIM_ERR_NOMEM);
    }
    std::vector<double> at(4 * (size_t)m);
    std::copy(axyz, axyz + 3 * m, at.begin());
    std::copy(radii, radii + m, at.begin() + 3 * m);
    if (hipMemcpy(d_at, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("hipMemcpy (atoms)");
        return fail(IBIM_ERR_HIP);
    }
    const double init = gp->init < 0 ? gp->init : -2.0 * std::max(L.d.nx, std::max(L.d.ny, L.d.nz)) * L.d.h;
    k_fill<<<nb, 256, 0, L.st>>>(L.g, L.n, init);
    k_to_ord<<<nb, 256, 0, L.st>>>(L.g, L.n);
    {
        const int gy = (int)std::min<int64_t>(m, 65535), gz = (int)((m + gy - 1) / gy);
        dim3 grid(64, gy, gz);
        k_expand<<<grid, 256, 0, L.st>>>((unsigned long long *)L.g, L.d, d_at, d_at + 3 * m, (int)m, gp->probe,
                                         gp->band, 0);
    }
    k_from_ord<<<nb, 256, 0, L.st>>>(L.g, L.n);
    if (hipGetLastError() != hipSuccess) {
        set_error("expand launch failed");
        return fail(IBIM_ERR_HIP);
    }
    const double tube = gp->thickness * L.d.h;
    // evolve inward by the probe radius (main.cpp:51: final time 1.4, speed s)
    if ((rc = ls_rk3(L, 0, gp->probe, 1.0, 0.2, tube, gp->thres, &S->info.evolve_steps, nullptr))) return fail(rc);
    // flip the sign, then reinitialise twice with phi0 the current level set (main.cpp:53-60)
    k_scale<<<nb, 256, 0, L.st>>>(L.g, L.n, -1.0);
    for (int pass = 0; pass < 2; ++pass) {
        if (hipMemcpyAsync(L.phi0, L.g, (size_t)L.n * sizeof(double), hipMemcpyDeviceToDevice, L.st) != hipSuccess) {
            set_error("hipMemcpy (phi0)");
            return fail(IBIM_ERR_HIP);
        }
        if ((rc = ls_rk3(L, 1, gp->reinit_time, 1.0, 0.5, tube, gp->thres, &S->info.reinit_steps[pass],
                         S->info.quality)))
            return fail(rc);
    }
    if ((rc = ls_surface(L, *gp, S))) return fail(rc);
    (void)hipFree(d_at);
    d_at = nullptr;
    free_grid(L);
    *out = S;
    return IBIM_OK;
}

int ibim_surface_from_grid(const ibim_geom_params *gp, const double *phi, ibim_surface **out) {
    if (!out || !phi) {
        set_error("bad argument to ibim_surface_from_grid");
        return IBIM_ERR_ARG;
    }
    *out = nullptr;
    int rc = check_geom(gp);
    if (rc) return rc;
    LsRun L;
    L.d = GridDims{gp->dims[0], gp->dims[1], gp->dims[2], gp->lo[0], gp->lo[1], gp->lo[2], gp->h, 1.0 / gp->h};
    IBIM_HIP_CHECK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
    if ((rc = alloc_grid(L, false))) {
        free_grid(L);
        return rc;
    }
    if (hipMemcpy(L.g, phi, (size_t)L.n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        free_grid(L);
        set_error("hipMemcpy (phi)");
        return IBIM_ERR_HIP;
    }
    ibim_surface *S = new ibim_surface();
    if ((rc = ls_surface(L, *gp, S))) {
        free_grid(L);
        ibim_surface_destroy(S);
        return rc;
    }
    free_grid(L);
    *out = S;
    return IBIM_OK;
}

int ibim_surface_get_info(const ibim_surface *s, ibim_surface_info *info) {
    if (!s || !info) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    *info = s->info;
    return IBIM_OK;
}

int ibim_surface_copy(const ibim_surface *s, double *xyz, double *normals, double *weights) {
    if (!s) {
        set_error("surface is NULL");
        return IBIM_ERR_ARG;
    }
    if (xyz) IBIM_HIP_CHECK(hipMemcpy(xyz, s->xyz, 3 * (size_t)s->n * sizeof(double), hipMemcpyDeviceToHost));
    if (normals) IBIM_HIP_CHECK(hipMemcpy(normals, s->nrm, 3 * (size_t)s->n * sizeof(double), hipMemcpyDeviceToHost));
    if (weights) IBIM_HIP_CHECK(hipMemcpy(weights, s->w, (size_t)s->n * sizeof(double), hipMemcpyDeviceToHost));
    return IBIM_OK;
}

int ibim_surface_device(const ibim_surface *s, const double **d_xyz, const double **d_normals,
                        const double **d_weights) {
    if (!s) {
        set_error("surface is NULL");
        return IBIM_ERR_ARG;
    }
    if (d_xyz) *d_xyz = s->xyz;
    if (d_normals) *d_normals = s->nrm;
    if (d_weights) *d_weights = s->w;
    return IBIM_OK;
}

}  // extern "C"
