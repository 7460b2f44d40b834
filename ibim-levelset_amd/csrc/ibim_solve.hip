// ibim_solve.hip -- the steps either side of the matvec, on the device:
// right-hand side (electric.cpp:395-413), target sets for the reaction
// potential at atom centres (polarizeMap, electric.cpp:442-485) and restarted
// GMRES with modified Gram-Schmidt and Givens rotations (gmres.cpp:10-164;
// the inner solves of the mixed-precision refinement keep an FP64 basis too,
// orthogonalised by classical Gram-Schmidt).
//
// GMRES keeps the Krylov basis, the Hessenberg matrix, the rotations and the
// residual estimate in HBM; the host only reads the residual estimate once
// per Arnoldi step to decide when to stop, as the reference does.
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <mutex>

#include "ibim_internal.h"

namespace ibim {

int kernel_fmm_device(ibim_plan *pl, int kid, const double *d_q, double *d_out);
int matvec_device(ibim_plan *pl, const double *d_x, double *d_y, bool mixed);

namespace {

inline int nblk(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------------ RHS ---
__global__ __launch_bounds__(256) void k_rhs(int n, const int *__restrict__ perm, const double *__restrict__ sx,
                                             const double *__restrict__ sy, const double *__restrict__ sz,
                                             const double *__restrict__ snx, const double *__restrict__ sny,
                                             const double *__restrict__ snz, int m, const double *__restrict__ axyz,
                                             const double *__restrict__ aq, double k1, double k2,
                                             double *__restrict__ b) {
    __shared__ double s_a[4][256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const double px = valid ? sx[i] : 0.0, py = valid ? sy[i] : 0.0, pz = valid ? sz[i] : 0.0;
    const double nx = valid ? snx[i] : 0.0, ny = valid ? sny[i] : 0.0, nz = valid ? snz[i] : 0.0;
    double b1 = 0.0, b2 = 0.0;
    for (int a0 = 0; a0 < m; a0 += 256) {
        const int cnt = min(256, m - a0);
        __syncthreads();
        if ((int)threadIdx.x < cnt) {
            const int a = a0 + threadIdx.x;
            s_a[0][threadIdx.x] = axyz[3 * (size_t)a];
            s_a[1][threadIdx.x] = axyz[3 * (size_t)a + 1];
            s_a[2][threadIdx.x] = axyz[3 * (size_t)a + 2];
            s_a[3][threadIdx.x] = aq[a];
        }
        __syncthreads();
        if (valid)
            for (int a = 0; a < cnt; ++a) {
                const double dx = px - s_a[0][a], dy = py - s_a[1][a], dz = pz - s_a[2][a];
                const double d = dx * dx + dy * dy + dz * dz;
                const double ir = 1.0 / sqrt(d);
                const double q = s_a[3][a];
                b1 += q * ir * k1;
                b2 -= q * ir * ir * ir * (nx * dx + ny * dy + nz * dz) * k2;
            }
    }
    if (valid) {
        const int o = perm[i];
        b[o] = b1;
        b[(size_t)n + o] = b2;
    }
}

// ------------------------------------------------------- target sets ---
__global__ void k_target_keys(const double *__restrict__ xyz, int n, double cx, double cy, double cz, double rx,
                              double ry, double rz, int maxlev, uint64_t *__restrict__ keys, int *__restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = xyz[3 * (size_t)i], py = xyz[3 * (size_t)i + 1], pz = xyz[3 * (size_t)i + 2];
    uint64_t key = 0;
    for (int l = 0; l < maxlev; ++l) {  // same comparisons as the sources (bbfmm.cpp:121-131)
        int xb = px < cx ? 0 : 1;
        int yb = py < cy ? 0 : 1;
        int zb = pz < cz ? 0 : 1;
        key = (key << 3) | (uint64_t)(4 * zb + 2 * yb + xb);
        cx = cx + ((double)xb - 0.5) * rx;
        cy = cy + ((double)yb - 0.5) * ry;
        cz = cz + ((double)zb - 0.5) * rz;
        rx = rx * 0.5;
        ry = ry * 0.5;
        rz = rz * 0.5;
    }
    keys[i] = key;
    idx[i] = i;
}

__global__ void k_gather3(const int *__restrict__ perm, int n, const double *__restrict__ xyz, double *x, double *y,
                          double *z) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t o = (size_t)perm[i];
    x[i] = xyz[3 * o];
    y[i] = xyz[3 * o + 1];
    z[i] = xyz[3 * o + 2];
}

// target range of every node from its key prefix
__global__ void k_target_ranges(int nnodes, const uint64_t *__restrict__ nkey, const int *__restrict__ level,
                                const int *__restrict__ flags, int maxlev, const uint64_t *__restrict__ tkeys, int m,
                                int *__restrict__ tbeg, int *__restrict__ tend, unsigned char *__restrict__ is_tleaf,
                                unsigned char *__restrict__ need, unsigned char *__restrict__ m2l) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nnodes) return;
    const int sh = 3 * (maxlev - level[id]);
    const uint64_t lo = nkey[id] << sh, hi = (nkey[id] + 1) << sh;
    int a = 0, b = m;
    while (a < b) {
        int mid = (a + b) >> 1;
        if (tkeys[mid] < lo) a = mid + 1;
        else b = mid;
    }
    int c = a, d = m;
    while (c < d) {
        int mid = (c + d) >> 1;
        if (tkeys[mid] < hi) c = mid + 1;
        else d = mid;
    }
    tbeg[id] = a;
    tend[id] = c;
    const bool has = c > a;
    is_tleaf[id] = (has && (flags[id] & kLeaf)) ? 1 : 0;
    need[id] = has ? 1 : 0;
    m2l[id] = (has && id != 0) ? 1 : 0;
}

__global__ void k_scatter_neg_f0(int m, const int *__restrict__ tperm, const double *__restrict__ F,
                                 double *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    out[tperm[i]] = -F[i];  // K2[qpsi] - sum K1a[qa] = -F0 (electric.cpp:482)
}

// -------------------------------------------------------------- GMRES ---
constexpr int kRB = 256;   // threads per reduction block
constexpr int kRG = 512;   // reduction blocks (fixed: deterministic sums)

__device__ __forceinline__ double block_sum(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kRB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

// sum of the kRG partials, same order in every block
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, double *sh) {
    double v = 0.0;
    for (int g = threadIdx.x; g < kRG; g += kRB) v += part[g];
    return block_sum(v, sh);
}

__global__ __launch_bounds__(kRB) void k_dot_part(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                  double *__restrict__ part) {
    __shared__ double sh[kRB];
    double v = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) v += a[i] * b[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// h = <w, v_k> from partials; H(k,i) = h; w -= h v_k; next partial of
// <w, v_{k+1}> (or <w, w> when last)
__global__ __launch_bounds__(kRB) void k_mgs(double *__restrict__ w, const double *__restrict__ vk,
                                             const double *__restrict__ vnext, int n,
                                             const double *__restrict__ part_in, double *__restrict__ part_out,
                                             double *__restrict__ hslot) {
    __shared__ double sh[kRB];
    const double h = sum_partials(part_in, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) *hslot = h;
    double v = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) {
        const double wi = w[i] - h * vk[i];
        w[i] = wi;
        v += wi * (vnext ? vnext[i] : wi);
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part_out[blockIdx.x] = v;
}

// H(i+1,i) = |w|; v_{i+1} = w * (1/|w|) (VT: the basis type)
template <typename VT>
__global__ __launch_bounds__(kRB) void k_normalize(const double *__restrict__ w, VT *__restrict__ vout, int n,
                                                   const double *__restrict__ part_in, double *__restrict__ hslot) {
    __shared__ double sh[kRB];
    const double nr = sqrt(sum_partials(part_in, sh));
    if (blockIdx.x == 0 && threadIdx.x == 0) *hslot = nr;
    const double inv = 1.0 / nr;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) vout[i] = (VT)(w[i] * inv);
}

// v0 = r * (1/beta); s = 0; s(0) = beta
template <typename VT>
__global__ __launch_bounds__(kRB) void k_start(const double *__restrict__ r, VT *__restrict__ v0, int n,
                                               const double *__restrict__ part_in, double *__restrict__ s, int m,
                                               double *__restrict__ beta_out) {
    __shared__ double sh[kRB];
    const double beta = sqrt(sum_partials(part_in, sh));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int q = 0; q <= m; ++q) s[q] = 0.0;
        s[0] = beta;
        *beta_out = beta;
    }
    const double inv = 1.0 / beta;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) v0[i] = (VT)(r[i] * inv);
}

// ---- inner (mixed-precision) Arnoldi steps: FP64 basis, classical
// Gram-Schmidt.  The inner solves of the refinement reduce the residual by
// about 1e-5 (what the FP32 / FP16 operator supports), and the FP64 outer
// residual decides convergence, so their FP64 basis is orthogonalised with
// one pass over it for all projections and one for the update, instead of
// two passes per basis vector (modified Gram-Schmidt).

// partials of <w, v_k>, k < kn: part[k kRG + block]
__global__ __launch_bounds__(kRB) void k_cgs_dots(const double *__restrict__ w, const double *__restrict__ V,
                                                  size_t ldv, int n, int kn, double *__restrict__ part) {
    __shared__ double sh[kRB];
    for (int k = 0; k < kn; ++k) {
        const double *vk = V + (size_t)k * ldv;
        double v = 0.0;
        for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) v += w[i] * vk[i];
        v = block_sum(v, sh);
        if (threadIdx.x == 0) part[(size_t)k * kRG + blockIdx.x] = v;
    }
}

// one block per k: coef[k] = <w, v_k>; H(k,i) = coef[k] (first pass) or
// += coef[k] (reorthogonalisation)
__global__ __launch_bounds__(kRB) void k_cgs_coef(const double *__restrict__ part, double *__restrict__ coef,
                                                  double *__restrict__ hcol, int first) {
    __shared__ double sh[kRB];
    const int k = blockIdx.x;
    const double c = sum_partials(part + (size_t)k * kRG, sh);
    if (threadIdx.x == 0) {
        coef[k] = c;
        hcol[k] = first ? c : hcol[k] + c;
    }
}

// w -= sum_k coef[k] v_k (k ascending); partial |w|^2
__global__ __launch_bounds__(kRB) void k_cgs_update(double *__restrict__ w, const double *__restrict__ V,
                                                    size_t ldv, int n, const double *__restrict__ coef, int kn,
                                                    double *__restrict__ part) {
    __shared__ double c[128];
    __shared__ double sh[kRB];
    for (int k = threadIdx.x; k < kn; k += kRB) c[k] = coef[k];
    __syncthreads();
    double v = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) {
        double wi = w[i];
        for (int k = 0; k < kn; ++k) wi -= c[k] * V[(size_t)k * ldv + i];
        w[i] = wi;
        v += wi * wi;
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__device__ void gen_rot(double dx, double dy, double &cs, double &sn) {  // gmres.cpp:23-36
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (fabs(dy) > fabs(dx)) {
        double temp = dx / dy;
        sn = 1.0 / sqrt(1.0 + temp * temp);
        cs = temp * sn;
    } else {
        double temp = dy / dx;
        cs = 1.0 / sqrt(1.0 + temp * temp);
        sn = temp * cs;
    }
}

__device__ void apply_rot(double &dx, double &dy, double cs, double sn) {  // gmres.cpp:39-44
    double temp = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = temp;
}

// gmres.cpp:121-130 on column i of H (column-major, leading dimension ld)
__global__ void k_givens(double *__restrict__ H, int ld, double *__restrict__ s, double *__restrict__ cs,
                         double *__restrict__ sn, int i, double normb, double *__restrict__ resid) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double *h = H + (size_t)i * ld;
    for (int k = 0; k < i; ++k) apply_rot(h[k], h[k + 1], cs[k], sn[k]);
    gen_rot(h[i], h[i + 1], cs[i], sn[i]);
    apply_rot(h[i], h[i + 1], cs[i], sn[i]);
    apply_rot(s[i], s[i + 1], cs[i], sn[i]);
    *resid = fabs(s[i + 1]) / normb;
}

// Update (gmres.cpp:10-21): back substitution, then x += sum_j y_j v_j
template <typename VT>
__global__ __launch_bounds__(kRB) void k_update(double *__restrict__ x, const VT *__restrict__ V, size_t ldv, int n,
                                                const double *__restrict__ H, int ld, const double *__restrict__ s,
                                                int k) {
    __shared__ double y[128];
    if (threadIdx.x == 0) {
        for (int i = 0; i <= k; ++i) y[i] = s[i];
        for (int i = k; i >= 0; i--) {
            y[i] /= H[(size_t)i * ld + i];
            for (int j = i - 1; j >= 0; j--) y[j] -= H[(size_t)i * ld + j] * y[i];
        }
    }
    __syncthreads();
    for (int e = blockIdx.x * kRB + threadIdx.x; e < n; e += gridDim.x * kRB) {
        double acc = x[e];
        for (int j = 0; j <= k; ++j) acc += y[j] * (double)V[(size_t)j * ldv + e];
        x[e] = acc;
    }
}

// r = b - p and partial |r|^2
__global__ __launch_bounds__(kRB) void k_residual(const double *__restrict__ b, const double *__restrict__ p,
                                                  double *__restrict__ r, int n, double *__restrict__ part) {
    __shared__ double sh[kRB];
    double v = 0.0;
    for (int i = blockIdx.x * kRB + threadIdx.x; i < n; i += kRG * kRB) {
        const double ri = b[i] - p[i];
        r[i] = ri;
        v += ri * ri;
    }
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_sqrt_sum(const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double sh[kRB];
    double v = sum_partials(part, sh);
    if (threadIdx.x == 0) *out = sqrt(v);
}

}  // namespace

int rhs_device(ibim_plan *pl, int m, const double *d_axyz, const double *d_aq, double *d_b) {
    const double pi = 3.14159265358979323846;
    const double dE = pl->prm.eps_out, dI = pl->prm.eps_in;
    const double k1 = 1.0 / (dI * 4.0 * pi * (0.5 * (1 + dE / dI)));
    const double k2 = 1.0 / (dI * 4.0 * pi * (0.5 * (1 + dI / dE)));
    StageScope sc(pl, kStRhs);
    k_rhs<<<nblk(pl->n), 256, 0, pl->stream>>>(pl->n, pl->perm, pl->sx, pl->sy, pl->sz, pl->snx, pl->sny, pl->snz, m,
                                               d_axyz, d_aq, k1, k2, d_b);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

int TmpAlloc::alloc(void **p, size_t bytes) {
    *p = nullptr;
    bytes = bytes ? (bytes + 255) & ~(size_t)255 : 256;
    if (arena) {
        if (bytes > cap - used) {
            set_error("target set needs more than the " + std::to_string(cap >> 20) + " MiB scratch arena (" +
                      std::to_string((used + bytes) >> 20) + " MiB asked so far)");
            return IBIM_ERR_ARG;
        }
        *p = arena + used;
        used += bytes;
        return IBIM_OK;
    }
    hipError_t e = dev_malloc(p, bytes);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc (targets): ") + hipGetErrorString(e));
        return IBIM_ERR_NOMEM;
    }
    owned.push_back(*p);
    return IBIM_OK;
}

void TmpAlloc::release() {
    for (void *q : owned) (void)hipFree(q);
    owned.clear();
    used = 0;
}

int make_target_set(ibim_plan *pl, int m, const double *d_txyz, TargetSet &ts, TmpAlloc &tmp, int **perm_out) {
    hipStream_t st = pl->stream;
    const int nn = pl->nnodes, maxlev = pl->prm.max_level;
    auto talloc = [&](auto **p, size_t cnt) -> int { return tmp.get(p, cnt); };
    uint64_t *k_in, *k_out;
    int *i_in, *perm, *tbeg, *tend, *leaves, *work, *d_num;
    double *x, *y, *z;
    unsigned char *fl_leaf, *need, *fl_m2l;
    int rc;
    if ((rc = talloc(&k_in, m)) || (rc = talloc(&k_out, m)) || (rc = talloc(&i_in, m)) || (rc = talloc(&perm, m)) ||
        (rc = talloc(&x, m)) || (rc = talloc(&y, m)) || (rc = talloc(&z, m)) || (rc = talloc(&tbeg, nn)) ||
        (rc = talloc(&tend, nn)) || (rc = talloc(&leaves, nn)) || (rc = talloc(&work, nn)) ||
        (rc = talloc(&fl_leaf, nn)) || (rc = talloc(&need, nn)) || (rc = talloc(&fl_m2l, nn)) ||
        (rc = talloc(&d_num, 2)))
        return rc;
    k_target_keys<<<nblk(m), 256, 0, st>>>(d_txyz, m, pl->root_c[0], pl->root_c[1], pl->root_c[2], pl->root_r[0],
                                           pl->root_r[1], pl->root_r[2], maxlev, k_in, i_in);
    IBIM_LAUNCH_CHECK();
    size_t bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, i_in, perm, m, 0,
                                                      std::max(1, 3 * maxlev), st));
    void *sort_tmp;
    if ((rc = talloc((char **)&sort_tmp, bytes))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, bytes, k_in, k_out, i_in, perm, m, 0,
                                                      std::max(1, 3 * maxlev), st));
    k_gather3<<<nblk(m), 256, 0, st>>>(perm, m, d_txyz, x, y, z);
    k_target_ranges<<<nblk(nn), 256, 0, st>>>(nn, pl->nkey, pl->level, pl->flags, maxlev, k_out, m, tbeg, tend,
                                              fl_leaf, need, fl_m2l);
    IBIM_LAUNCH_CHECK();
    hipcub::CountingInputIterator<int> it(0);
    size_t b1 = 0, b2 = 0;
    IBIM_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, b1, it, fl_leaf, leaves, d_num, nn, st));
    IBIM_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, b2, it, fl_m2l, work, d_num + 1, nn, st));
    void *sel_tmp;
    if ((rc = talloc((char **)&sel_tmp, std::max(b1, b2)))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceSelect::Flagged(sel_tmp, b1, it, fl_leaf, leaves, d_num, nn, st));
    IBIM_HIP_CHECK(hipcub::DeviceSelect::Flagged(sel_tmp, b2, it, fl_m2l, work, d_num + 1, nn, st));
    int h_num[2];
    IBIM_HIP_CHECK(hipMemcpyAsync(h_num, d_num, sizeof(h_num), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    ts.n = m;
    ts.x = x;
    ts.y = y;
    ts.z = z;
    ts.tbeg = tbeg;
    ts.tend = tend;
    ts.leaves = leaves;
    ts.nleaves = h_num[0];
    ts.m2l_work = work;
    ts.n_m2l_work = h_num[1];
    ts.need_local = need;
    *perm_out = perm;
    return IBIM_OK;
}

int single_charge_device(ibim_plan *pl, int kid, const double *d_q);
int scatter_field_device(ibim_plan *pl, int kid, const int *perm, const double *F, int nt, double *d_out);

// polarizeMap (kid < 0) or one kernel FMM (kid = 0..15) at m target points.
// With an arena (`tmp.arena` set) every temporary comes from it and the far
// field runs without the V-pair GEMM and its pair tables (V and X lists
// evaluated directly) and the near field without work units (one block per
// target leaf): nothing is allocated.
int target_fmm_device(ibim_plan *pl, int kid, int m, const double *d_axyz, const double *d_in, double *d_out,
                      TmpAlloc &tmp) {
    StageScope sc(pl, kStPolarize);
    const bool lean = tmp.arena != nullptr;
    TargetSet ts;
    int *tperm = nullptr;
    double *F = nullptr;
    int rc = make_target_set(pl, m, d_axyz, ts, tmp, &tperm);
    if (rc == IBIM_OK) rc = tmp.get(&F, 4 * (size_t)m);
    VPairs vp;
    if (rc == IBIM_OK && pl->ops && !lean) {
        rc = build_vpairs(pl, ts.need_local, vp, &tmp.owned);
        ts.vpairs = &vp;
    }
    P2PWork work;
    if (rc == IBIM_OK && !lean) rc = build_p2p_ranges(pl, ts, work, &tmp.owned);
    if (rc == IBIM_OK) rc = kid < 0 ? form_charges(pl, d_in) : single_charge_device(pl, kid, d_in);
    if (rc == IBIM_OK) rc = fmm_forward(pl, ts, F, m);
    if (rc == IBIM_OK) {
        if (kid < 0) k_scatter_neg_f0<<<nblk(m), 256, 0, pl->stream>>>(m, tperm, F, d_out);
        else rc = scatter_field_device(pl, kid, tperm, F, m, d_out);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
        if (e != hipSuccess) {
            set_error(std::string("target FMM: ") + hipGetErrorString(e));
            rc = IBIM_ERR_HIP;
        }
    } else {
        (void)hipStreamSynchronize(pl->stream);
    }
    tmp.release();
    return rc;
}

int target_fmm_device(ibim_plan *pl, int kid, int m, const double *d_axyz, const double *d_in, double *d_out) {
    TmpAlloc tmp;
    return target_fmm_device(pl, kid, m, d_axyz, d_in, d_out, tmp);
}

int polarize_device(ibim_plan *pl, int m, const double *d_axyz, const double *d_x, double *d_out) {
    return target_fmm_device(pl, -1, m, d_axyz, d_x, d_out);
}

// ---------------------------------------------- polarize without allocating --
// The scratch of ibim_polarize_device_noalloc: one array per GPU (each device
// holds its own copy of this file-scope array from the library's load to the
// end of the process).  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device, kept per device in
// g_scratch; the array's name on the host is not a device address.
constexpr size_t kNoallocScratch = (size_t)64 << 20;  // bytes
static __device__ __align__(256) unsigned char scratch[kNoallocScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];   // one call at a time per device (the call is synchronous)
void *g_scratch[kMaxDevices] = {};      // device address of `scratch`, per device
}  // namespace

// the target set's per-call scratch in bytes, an upper bound from m targets
// and the plan's nodes: keys, indices, coordinates and fields per target,
// ranges, work lists and flags per node, and the radix sort's and
// selection's temporaries
static size_t noalloc_bytes(int m, int nn) {
    return 256 * 24 + (size_t)m * (8 + 8 + 4 + 4 + 3 * 8 + 4 * 8) + (size_t)nn * (4 * 4 + 3) +
           (size_t)m * 24 + (size_t)nn * 8 + ((size_t)1 << 20);
}

int polarize_noalloc(ibim_plan *pl, int m, const double *d_axyz, const double *d_x, double *d_out) {
    int dev = 0;
    IBIM_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices || dev != pl->device) {
        set_error("polarize (no-alloc): the plan's device must be current (and below 64)");
        return IBIM_ERR_ARG;
    }
    const size_t need = noalloc_bytes(m, pl->nnodes);
    if (need > kNoallocScratch) {
        set_error("polarize (no-alloc): " + std::to_string(m) + " targets on a plan of " +
                  std::to_string(pl->nnodes) + " boxes need about " + std::to_string(need >> 20) +
                  " MiB of scratch; the maximum is " + std::to_string(kNoallocScratch >> 20) + " MiB");
        return IBIM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu[dev]);
    if (!g_scratch[dev]) {  // once per device, with that device current
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress (scratch): ") + hipGetErrorString(e));
            return IBIM_ERR_HIP;
        }
        g_scratch[dev] = p;
    }
    TmpAlloc tmp;
    tmp.arena = (char *)g_scratch[dev];
    tmp.cap = kNoallocScratch;
    return target_fmm_device(pl, -1, m, d_axyz, d_x, d_out, tmp);
}

constexpr int kHostScalars = 8;

// the workspace for GMRES(restart): the FP64 basis, and (cgs) the
// projection partials of classical Gram-Schmidt for the inner solves,
// allocated when first needed
static int gmres_alloc(ibim_plan *pl, int restart, bool cgs, bool basis = true) {
    const size_t n2 = 2 * (size_t)pl->n;
    int rc;
    // earlier buffers stay on the plan's allocation list and are freed with it
    if (!pl->gm_w || pl->gm_restart < restart) {
        if ((rc = dalloc(pl, &pl->gm_w, n2)) || (rc = dalloc(pl, &pl->gm_r, n2)) || (rc = dalloc(pl, &pl->gm_x, n2)) ||
            (rc = dalloc(pl, &pl->gm_part, 4 * (size_t)kRG)) ||
            (rc = dalloc(pl, &pl->gm_small, (size_t)(restart + 1) * (restart + 1) + 4 * (size_t)(restart + 1) + 8)))
            return rc;
        pl->gm_V = nullptr;
        pl->gm_hpart = nullptr;
        pl->gm_restart = restart;
    }
    const int rs = pl->gm_restart;
    if (basis && !pl->gm_V && (rc = dalloc(pl, &pl->gm_V, (size_t)(rs + 1) * n2))) return rc;
    if (basis && cgs && !pl->gm_hpart &&
        ((rc = dalloc(pl, &pl->gm_hpart, (size_t)(rs + 1) * kRG)) || (rc = dalloc(pl, &pl->gm_coef, (size_t)rs + 1))))
        return rc;
    if (!pl->gm_host) IBIM_HIP_CHECK(hipHostMalloc((void **)&pl->gm_host, kHostScalars * sizeof(double)));
    return IBIM_OK;
}

// Device scalar -> slot k of the pinned host buffer (a direct copy: no
// staging through a pageable bounce buffer); read the slot after the stream
// synchronises.
static inline hipError_t fetch_scalar(ibim_plan *pl, const double *d, int k) {
    return hipMemcpyAsync(pl->gm_host + k, d, sizeof(double), hipMemcpyDeviceToHost, pl->stream);
}

// restarted GMRES, gmres.cpp:50-164, all vectors in HBM; `mixed` selects
// the mixed-precision matvec for every operator application
static int gmres_core(ibim_plan *pl, const double *d_b, double *d_x, int restart, int maxit, double tol,
                      ibim_solve_info *info, bool mixed) {
    if (restart < 1 || restart > 120 || maxit < 0) {
        set_error("gmres: restart must be in 1..120 and maxit >= 0");
        return IBIM_ERR_ARG;
    }
    // the inner solves of the refinement orthogonalise by classical
    // Gram-Schmidt; the FP64 solve by the reference's modified Gram-Schmidt
    const bool cgs = mixed;
    int rc = gmres_alloc(pl, restart, cgs);
    if (rc) return rc;
    hipStream_t st = pl->stream;
    const int m = restart;
    const int n2 = 2 * pl->n;
    const size_t ldv = (size_t)n2;
    const int ld = m + 1;
    double *V = pl->gm_V, *w = pl->gm_w, *r = pl->gm_r, *p = pl->gm_x;
    double *partA = pl->gm_part, *partB = pl->gm_part + kRG;
    double *H = pl->gm_small;
    double *s = H + (size_t)ld * ld, *cs = s + ld, *sn = cs + ld, *scal = sn + ld;  // scal: beta, resid, normb
    IBIM_HIP_CHECK(hipMemsetAsync(H, 0, (size_t)ld * ld * sizeof(double), st));

    int nmv = 0, inner = 0;
    // normb
    k_dot_part<<<kRG, kRB, 0, st>>>(d_b, d_b, n2, partA);
    k_sqrt_sum<<<1, kRB, 0, st>>>(partA, scal + 2);
    IBIM_LAUNCH_CHECK();
    IBIM_HIP_CHECK(fetch_scalar(pl, scal + 2, 2));
    // r = b - A x; for x = 0 this is b exactly (the matvec of 0 is 0), so the
    // operator is not applied
    k_dot_part<<<kRG, kRB, 0, st>>>(d_x, d_x, n2, partB);
    k_sqrt_sum<<<1, kRB, 0, st>>>(partB, scal + 3);
    IBIM_LAUNCH_CHECK();
    IBIM_HIP_CHECK(fetch_scalar(pl, scal + 3, 3));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    double normb = pl->gm_host[2];
    const double xnorm = pl->gm_host[3];
    if (xnorm != 0.0) {
        if ((rc = matvec_device(pl, d_x, p, mixed))) return rc;
        nmv++;
    } else {
        IBIM_HIP_CHECK(hipMemsetAsync(p, 0, (size_t)n2 * sizeof(double), st));
    }
    k_residual<<<kRG, kRB, 0, st>>>(d_b, p, r, n2, partA);
    k_sqrt_sum<<<1, kRB, 0, st>>>(partA, scal);
    IBIM_LAUNCH_CHECK();
    IBIM_HIP_CHECK(fetch_scalar(pl, scal, 0));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    double beta = pl->gm_host[0];
    if (normb == 0.0) normb = 1;
    double res = beta / normb;
    int ret = IBIM_NOT_CONVERGED;
    // the reference's iteration table (gmres.cpp:79-106): iteration, the
    // relative residual, the seconds the iteration took
    auto clock_now = [] { return std::chrono::steady_clock::now(); };
    auto t_last = clock_now();
    if (pl->trace) {
        std::printf("=============== GMRES =================\n    iter    |  rel error   |   time   \n");
        std::fflush(stdout);
    }
    auto trace_line = [&](int it, double r) {
        if (!pl->trace) return;
        const auto t = clock_now();
        std::printf("%6d%20.6e%12.6f\n", it, r, std::chrono::duration<double>(t - t_last).count());
        std::fflush(stdout);
        t_last = t;
    };
    if (res <= tol) {
        ret = IBIM_OK;
    } else {
        int i = 0, j = 1;
        bool done = false;
        while (j <= maxit && !done) {
            // v0 = r / beta (partials of |r|^2 are in partA)
            k_start<double><<<kRG, kRB, 0, st>>>(r, V, n2, partA, s, m, scal);
            IBIM_LAUNCH_CHECK();
            for (i = 0; i < m && j <= maxit; i++, j++) {
                const double *vi = V + (size_t)i * ldv;
                if ((rc = matvec_device(pl, vi, w, mixed))) return rc;
                nmv++;
                inner++;
                double *hcol = H + (size_t)i * ld;
                {
                    StageScope sc(pl, kStOrth);
                    if (cgs) {
                        k_cgs_dots<<<kRG, kRB, 0, st>>>(w, V, ldv, n2, i + 1, pl->gm_hpart);
                        k_cgs_coef<<<i + 1, kRB, 0, st>>>(pl->gm_hpart, pl->gm_coef, hcol, 1);
                        k_cgs_update<<<kRG, kRB, 0, st>>>(w, V, ldv, n2, pl->gm_coef, i + 1, partA);
                        k_normalize<double><<<kRG, kRB, 0, st>>>(w, V + (size_t)(i + 1) * ldv, n2, partA, hcol + i + 1);
                    } else {
                        k_dot_part<<<kRG, kRB, 0, st>>>(w, V, n2, partA);  // <w, v_0>
                        double *pin = partA, *pout = partB;
                        for (int k = 0; k <= i; ++k) {
                            const double *vnext = (k < i) ? V + (size_t)(k + 1) * ldv : nullptr;
                            k_mgs<<<kRG, kRB, 0, st>>>(w, V + (size_t)k * ldv, vnext, n2, pin, pout, hcol + k);
                            std::swap(pin, pout);
                        }
                        k_normalize<double><<<kRG, kRB, 0, st>>>(w, V + (size_t)(i + 1) * ldv, n2, pin, hcol + i + 1);
                    }
                    k_givens<<<1, 64, 0, st>>>(H, ld, s, cs, sn, i, normb, scal + 1);
                }
                IBIM_LAUNCH_CHECK();
                IBIM_HIP_CHECK(fetch_scalar(pl, scal + 1, 1));
                IBIM_HIP_CHECK(hipStreamSynchronize(st));
                res = pl->gm_host[1];
                trace_line(j, res);
                if (res < tol) {
                    k_update<double><<<kRG, kRB, 0, st>>>(d_x, V, ldv, n2, H, ld, s, i);
                    IBIM_LAUNCH_CHECK();
                    ret = IBIM_OK;
                    done = true;
                    break;
                }
            }
            if (done) break;
            k_update<double><<<kRG, kRB, 0, st>>>(d_x, V, ldv, n2, H, ld, s, i - 1);
            IBIM_LAUNCH_CHECK();
            if ((rc = matvec_device(pl, d_x, p, mixed))) return rc;
            nmv++;
            k_residual<<<kRG, kRB, 0, st>>>(d_b, p, r, n2, partA);
            k_sqrt_sum<<<1, kRB, 0, st>>>(partA, scal);
            IBIM_LAUNCH_CHECK();
            IBIM_HIP_CHECK(fetch_scalar(pl, scal, 0));
            IBIM_HIP_CHECK(hipStreamSynchronize(st));
            beta = pl->gm_host[0];
            if ((res = beta / normb) < tol) {
                ret = IBIM_OK;
                break;
            }
        }
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    if (pl->trace) {
        std::printf(ret == IBIM_OK ? "=============== CONVERGED =============\n"
                                   : "============= NOT CONVERGED ===========\n");
        std::fflush(stdout);
    }
    if (info) {
        info->iterations = inner;
        info->matvecs = nmv;
        info->rel_residual = res;
        info->converged = ret == IBIM_OK ? 1 : 0;
        info->outer_steps = 0;
    }
    return ret;
}

}  // namespace ibim

namespace ibim {
namespace {
__global__ void k_axpy1(double *__restrict__ x, const double *__restrict__ d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// the plan's sorted points back in the caller's order (perm[sorted] = original)
__global__ void k_unsort_points(const int *__restrict__ perm, int n, const double *__restrict__ sx,
                                const double *__restrict__ sy, const double *__restrict__ sz,
                                const double *__restrict__ snx, const double *__restrict__ sny,
                                const double *__restrict__ snz, const double *__restrict__ sw, double *__restrict__ xyz,
                                double *__restrict__ nrm, double *__restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = (size_t)perm[i];
    xyz[3 * o] = sx[i];
    xyz[3 * o + 1] = sy[i];
    xyz[3 * o + 2] = sz[i];
    nrm[3 * o] = snx[i];
    nrm[3 * o + 1] = sny[i];
    nrm[3 * o + 2] = snz[i];
    w[o] = sw[i];
}
}  // namespace

// The plan whose matvec the inner solves use when prm.inner_order is set:
// the same points, leaf capacity and kernels at Chebyshev order inner_order,
// in this plan's mixed form, built on first use (its points are this plan's,
// put back in the caller's order) and kept with this plan.  Its work is
// ordered on this plan's stream, and it records the stages this plan does.
static int inner_plan(ibim_plan *pl, ibim_plan **out) {
    if (!pl->inner) {
        const int n = pl->n;
        double *xyz = nullptr, *nrm = nullptr, *w = nullptr;
        int rc = IBIM_OK;
        if (dev_malloc(&xyz, 3 * (size_t)n * sizeof(double)) != hipSuccess ||
            dev_malloc(&nrm, 3 * (size_t)n * sizeof(double)) != hipSuccess ||
            dev_malloc(&w, (size_t)n * sizeof(double)) != hipSuccess) {
            set_error("hipMalloc (inner plan points)");
            rc = IBIM_ERR_NOMEM;
        }
        if (rc == IBIM_OK) {
            k_unsort_points<<<(n + 255) / 256, 256, 0, pl->stream>>>(pl->perm, n, pl->sx, pl->sy, pl->sz, pl->snx,
                                                                      pl->sny, pl->snz, pl->sw, xyz, nrm, w);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
            if (e != hipSuccess) {
                set_error(std::string("inner plan points: ") + hipGetErrorString(e));
                rc = IBIM_ERR_HIP;
            }
        }
        if (rc == IBIM_OK) {
            ibim_params q = pl->prm;
            q.p = pl->prm.inner_order;
            q.inner_order = 0;
            rc = ibim_plan_create_device(&q, n, xyz, nrm, w, &pl->inner);  // (synchronises its build)
        }
        for (double *p : {xyz, nrm, w})
            if (p) (void)hipFree(p);
        if (rc) return rc;
        pl->inner->trace = pl->trace;
        pl->inner->timer.on = pl->timer.on;
        pl->inner->timer.mask = pl->timer.mask;
    }
    int rc = ibim_plan_set_stream(pl->inner, pl->stream);
    if (rc) return rc;
    *out = pl->inner;
    return IBIM_OK;
}

// Mixed precision: iterative refinement.  The residual r = b - A x is formed
// with the FP64 matvec; each correction solves A d = r with GMRES on the
// mixed-precision matvec (FP32 near and far field, FP64 Krylov basis,
// Hessenberg, rotations) to a relative tolerance no tighter than what the
// FP32 operator can deliver, and x += d.  Converged when |b - A x| / |b| <
// tol with the FP64 operator, so the answer is that of the FP64 solve.
static int gmres_refine(ibim_plan *pl, const double *d_b, double *d_x, int restart, int maxit, double tol,
                        ibim_solve_info *info) {
    hipStream_t st = pl->stream;
    const int n2 = 2 * pl->n;
    int rc;
    if (!pl->gm_rr && ((rc = dalloc(pl, &pl->gm_rr, (size_t)n2)) || (rc = dalloc(pl, &pl->gm_d, (size_t)n2))))
        return rc;
    // the inner solves' operator: this plan's mixed matvec, or that of its
    // lower-order companion (the FP64 residual below stays this plan's)
    ibim_plan *ip = pl;
    if (pl->prm.inner_order > 0 && (rc = inner_plan(pl, &ip))) {
        if (rc != IBIM_ERR_NOMEM) return rc;
        ip = pl;  // no room for the companion plan: the inner solves use this plan's own operator
    }
    if ((rc = gmres_alloc(pl, restart, true, ip == pl))) return rc;
    double *scal = pl->gm_small;  // borrowed scratch for the two norms below (before any inner solve)
    double *part = pl->gm_part + 3 * kRG;
    auto norm = [&](const double *v, double *out) -> int {
        k_dot_part<<<kRG, kRB, 0, st>>>(v, v, n2, part);
        k_sqrt_sum<<<1, kRB, 0, st>>>(part, scal);
        IBIM_LAUNCH_CHECK();
        IBIM_HIP_CHECK(fetch_scalar(pl, scal, 4));
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
        *out = pl->gm_host[4];
        return IBIM_OK;
    };
    double normb = 0.0, rn = 0.0;
    if ((rc = norm(d_b, &normb))) return rc;
    if (normb == 0.0) normb = 1.0;
    int nmv = 0, inner = 0, outer = 0;
    double xnorm = 0.0;
    if ((rc = norm(d_x, &xnorm))) return rc;
    bool x_zero = xnorm == 0.0;
    auto residual = [&]() -> int {  // gm_rr = b - A64 x, rn = |gm_rr|
        if (x_zero) {  // A 0 = 0 exactly
            IBIM_HIP_CHECK(hipMemsetAsync(pl->gm_x, 0, (size_t)n2 * sizeof(double), st));
            x_zero = false;
        } else {
            int r2 = matvec_device(pl, d_x, pl->gm_x, false);
            if (r2) return r2;
            ++nmv;
        }
        k_residual<<<kRG, kRB, 0, st>>>(d_b, pl->gm_x, pl->gm_rr, n2, part);
        k_sqrt_sum<<<1, kRB, 0, st>>>(part, scal);
        IBIM_LAUNCH_CHECK();
        IBIM_HIP_CHECK(fetch_scalar(pl, scal, 5));
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
        rn = pl->gm_host[5];
        return IBIM_OK;
    };
    if ((rc = residual())) return rc;
    // relative residual the inner solve aims at: the FP32 matvec supports 1e-6;
    // the FP16 V-list (per-box scales make it linear only to ~1e-5) 1e-5.  An
    // operator of lower order differs from the order-p one by ~1e-4, and a
    // step with it lowers the FP64 residual by only ~5e-4 to 1e-3 however far
    // the inner solve goes (the correction's error is the operators'
    // difference): 1e-3, since iterations beyond that buy nothing
    auto floor_of = [&](const ibim_plan *q) {
        if (q != pl) return 1e-3;
        return pl->prm.precision == IBIM_PREC_MIXED_F16 ? 1e-5 : 1e-6;
    };
    double floor_mixed = floor_of(ip);
    while (rn / normb >= tol && inner < maxit && outer < 50) {
        const double inner_tol = std::max(floor_mixed, 0.5 * tol * normb / rn);
        IBIM_HIP_CHECK(hipMemsetAsync(pl->gm_d, 0, (size_t)n2 * sizeof(double), st));
        ibim_solve_info ii{};
        rc = gmres_core(ip, pl->gm_rr, pl->gm_d, restart, maxit - inner, inner_tol, &ii, true);
        if (rc < 0) return rc;
        inner += ii.iterations;
        nmv += ii.matvecs;
        k_axpy1<<<(n2 + 255) / 256, 256, 0, st>>>(d_x, pl->gm_d, n2);
        IBIM_LAUNCH_CHECK();
        ++outer;
        const double before = rn;
        if ((rc = residual())) return rc;
        if (pl->trace) {
            std::printf("refinement step %d: %d inner iterations (order %d), FP64 relative residual %.6e\n", outer,
                        ii.iterations, ip->p, rn / normb);
            std::fflush(stdout);
        }
        if (ip != pl && !(rn < std::max(0.1, 10.0 * inner_tol) * before)) {
            // the FP64 residual fell by less than a tenth of what the inner
            // solve reached with the lower-order operator (or by less than
            // tenfold): the remaining steps use this plan's own operator
            ip = pl;
            floor_mixed = floor_of(ip);
            continue;
        }
        if (!(rn < before)) break;  // no progress: stop instead of looping
    }
    if (info) {
        info->iterations = inner;
        info->matvecs = nmv;
        info->rel_residual = rn / normb;
        info->converged = rn / normb < tol ? 1 : 0;
        info->outer_steps = outer;
    }
    return rn / normb < tol ? IBIM_OK : IBIM_NOT_CONVERGED;
}

int gmres_device(ibim_plan *pl, const double *d_b, double *d_x, int restart, int maxit, double tol,
                 ibim_solve_info *info) {
    if (pl->prm.precision != IBIM_PREC_FP64) return gmres_refine(pl, d_b, d_x, restart, maxit, tol, info);
    return gmres_core(pl, d_b, d_x, restart, maxit, tol, info, false);
}
}  // namespace ibim
