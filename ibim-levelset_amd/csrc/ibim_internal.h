// ibim_internal.h -- plan layout and device helpers shared by the HIP sources.
//
// Data layout in HBM (one plan = one surface point set):
//   points      sorted by octree key (stable in the original index), SoA
//               x, y, z, nx, ny, nz, w; perm[sorted] = original index
//   nodes       breadth first; the 8 children of a split node consecutive;
//               centre (3 doubles), level, parent, slot, flags, source range
//               [sbeg, send) in sorted order (a subtree is one range)
//   lists       CSR per kind (U, V, W, X) over all nodes
//   expansions  M (multipole) and L (local): 4 fields x r doubles per node,
//               fields = (qx, qy, qz, qpsi) upward, (F0, Fx, Fy, Fz) downward
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ibim.h"

namespace ibim {

constexpr double kEps = 1e-12;  // adjacency / matching tolerance, utils.h:40
constexpr int kMaxP = 8;
constexpr int kLeaf = 1;
constexpr int kEmpty = 2;

void set_error(const std::string &msg);

#define IBIM_HIP_CHECK(expr)                                                              \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ibim::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));           \
            return IBIM_ERR_HIP;                                                          \
        }                                                                                 \
    } while (0)

#define IBIM_LAUNCH_CHECK()                                                               \
    do {                                                                                  \
        hipError_t e_ = hipGetLastError();                                                \
        if (e_ != hipSuccess) {                                                           \
            ibim::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));      \
            return IBIM_ERR_HIP;                                                          \
        }                                                                                 \
    } while (0)

// Chebyshev tables, passed by value to kernels (bbfmm.h:273-367)
struct ChebTab {
    double xbar[kMaxP];            // standard nodes -cos((i+1/2) pi / p)
    double Tbar[kMaxP][kMaxP];     // Tbar[m][k] = T_k(xbar_m)
    double tr[2][kMaxP][kMaxP];    // tr[b][i][l] = S_p(child-b node i, xbar_l)
    double invp;                   // 1/p
};

// kernel constants of electric.cpp:52-58
struct PBConst {
    double kappa;
    double eps;       // eps_out / eps_in
    double inv_eps;   // eps_in / eps_out
    double vr;        // vacant radius tau*h
    double k2reg;     // K2 below the vacant radius (disk average), 0 if kappa == 0
    double inv4pi;
    int mode;         // IBIM_KERNEL_*
};

// 1/x for x > 0 normal: the hardware reciprocal refined by two Newton steps
// (within an ulp of the quotient; no special-value handling needed here)
__device__ __forceinline__ double recip(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}
__device__ __forceinline__ float recip(float x) { return 1.0f / x; }

// --------------------------------------------------------------------------
// The 16 kernels fused: one source s with weighted charges (qx, qy, qz, qp)
// acting on one target t, d = s - t.  Accumulates
//   F0 += sum_a K1a qa - K2 qp
//   Fa += sum_b K3ab qb - K4a qp
// which is exactly what dmapping forms from the 16 separate FMM results
// (electric.cpp:369-384); the kernels are electric.cpp:67-291:
//   K1a = d_a (eps e (t+1) - 1) / (4 pi r^3)          t = kappa r, e = exp(-t)
//   K2  = (1 - e) / (4 pi r)
//   K3ab = delta_ab (1 - e (t+1)) / r^3 + (e (t^2 + 3t + 3) - 3) d_a d_b / r^5
//   K4a = -d_a (e (t+1) / eps - 1) / (4 pi r^3)
// and below the vacant radius K1 = K3 = K4 = 0, K2 = k2reg.  pair_accum
// dispatches on the kernel family at run time, pair_accum_m at compile time.

// polynomial surrogate kernels (IBIM_KERNEL_POLY): K1a = d_a,
// K2 = 0.5 + 0.25 r^2, K3ab = delta_ab + 0.5 d_a d_b, K4a = -0.75 d_a
template <typename T>
__device__ __forceinline__ void pair_accum_poly(T dx, T dy, T dz, T qx, T qy, T qz, T qp, T &f0, T &fx, T &fy,
                                                T &fz) {
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T dq = dx * qx + dy * qy + dz * qz;
    f0 += dq - (T(0.5) + T(0.25) * r2) * qp;
    const T s = T(0.5) * dq + T(0.75) * qp;
    fx += qx + dx * s;
    fy += qy + dy * s;
    fz += qz + dz * s;
}

// the PB kernels.  Below the vacant radius only K2 = k2reg acts; that case
// is evaluated without a branch (the kernel terms are computed at a safe r
// and discarded).
template <typename T>
__device__ __forceinline__ void pair_accum_pb(T dx, T dy, T dz, T qx, T qy, T qz, T qp, const PBConst &k, T &f0,
                                              T &fx, T &fy, T &fz) {
    T r2 = dx * dx + dy * dy + dz * dz;
    T r = sqrt(r2);
    const bool reg = r < T(k.vr);
    if (reg) r = T(1);
    T ir = recip(r);
    T ir2 = ir * ir;
    T ir3 = ir2 * ir;
    T t = T(k.kappa) * r;
    T e = exp(-t);
    T et1 = e * (t + T(1));
    T c4 = T(k.inv4pi) * ir3;
    T g1 = (T(k.eps) * et1 - T(1)) * c4;
    T k2 = (T(1) - e) * ir * T(k.inv4pi);
    T a3 = (T(1) - et1) * ir3;
    T b3 = (e * ((t + T(3)) * t + T(3)) - T(3)) * ir3 * ir2;
    T g4 = (T(k.inv_eps) * et1 - T(1)) * c4;
    T dq = dx * qx + dy * qy + dz * qz;
    if (reg) {
        g1 = T(0);
        k2 = T(k.k2reg);
        a3 = T(0);
        b3 = T(0);
        g4 = T(0);
    }
    f0 += g1 * dq - k2 * qp;
    T s = b3 * dq + g4 * qp;
    fx += a3 * qx + dx * s;
    fy += a3 * qy + dy * s;
    fz += a3 * qz + dz * s;
}

// FP32 form for the mixed-precision near field and X list: the caller rounds
// an FP64-formed coordinate difference; reciprocal square root and exp use
// the hardware FP32 units.  At r = 0 or below the vacant radius 1/r and r
// are replaced by 0, which zeroes every term but K2.
template <>
__device__ __forceinline__ void pair_accum_pb<float>(float dx, float dy, float dz, float qx, float qy, float qz,
                                                     float qp, const PBConst &k, float &f0, float &fx, float &fy,
                                                     float &fz) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float ir0 = rsqrtf(r2);
    const float r0 = r2 * ir0;
    const bool reg = !(r0 >= (float)k.vr);  // also catches r2 == 0
    const float ir = reg ? 0.0f : ir0;
    const float r = reg ? 0.0f : r0;
    const float ir2 = ir * ir, ir3 = ir2 * ir;
    const float t = (float)k.kappa * r;
    const float e = __expf(-t);
    const float et1 = e * (t + 1.0f);
    const float c4 = (float)k.inv4pi * ir3;
    const float g1 = ((float)k.eps * et1 - 1.0f) * c4;
    const float k2 = reg ? (float)k.k2reg : (1.0f - e) * ir * (float)k.inv4pi;
    const float a3 = (1.0f - et1) * ir3;
    const float b3 = (e * ((t + 3.0f) * t + 3.0f) - 3.0f) * ir3 * ir2;
    const float g4 = ((float)k.inv_eps * et1 - 1.0f) * c4;
    const float dq = dx * qx + dy * qy + dz * qz;
    f0 += g1 * dq - k2 * qp;
    const float s = b3 * dq + g4 * qp;
    fx += a3 * qx + dx * s;
    fy += a3 * qy + dy * s;
    fz += a3 * qz + dz * s;
}

template <typename T>
__device__ __forceinline__ void pair_accum(T dx, T dy, T dz, T qx, T qy, T qz, T qp, const PBConst &k, T &f0,
                                           T &fx, T &fy, T &fz) {
    if (k.mode == IBIM_KERNEL_POLY) pair_accum_poly<T>(dx, dy, dz, qx, qy, qz, qp, f0, fx, fy, fz);
    else pair_accum_pb<T>(dx, dy, dz, qx, qy, qz, qp, k, f0, fx, fy, fz);
}

// one pair with the kernel family fixed at compile time (hot loops)
// FP32 near-field form of the PB kernels with 1/(4 pi) taken out of the
// sums.  Per source the caller supplies qs = qp / (4 pi) besides (qx, qy,
// qz); the pair adds
//   f0a += g1 dq,   f0b += k2 qs        so that F0 = f0a / (4 pi) - f0b
//   fa  += a3 qa + da (b3 dq + g4 qs)
// with g1 = (eps e (t+1) - 1) / r^3, g4 = (e (t+1) / eps - 1) / r^3, k2 =
// (1 - e) / r (4 pi k2reg below the vacant radius), a3 = (1 - e (t+1)) / r^3
// and b3 = (e (t^2 + 3t + 3) - 3) / r^5 = (e (t+1)(t+2) + e - 3) / r^5.
// The hardware rsq and exp2 are used without range fix-ups: r^2 = 0 and r
// below the vacant radius (r^2 < vr^2) are masked (1/r, hence r = r^2 / r,
// replaced by 0), and exp2 of the non-positive argument -kappa r log2(e) can
// only underflow to 0.
struct PBConst32 {
    float kappa, ck, eps, inv_eps, vr2, k2reg4pi;  // ck = -kappa log2(e), vr2 = vacant radius squared
};
__host__ __device__ inline PBConst32 pb_const32(const PBConst &k) {
    const double pi4 = 1.0 / k.inv4pi;
    return PBConst32{(float)k.kappa, (float)(-k.kappa * 1.4426950408889634), (float)k.eps, (float)k.inv_eps,
                     (float)(k.vr * k.vr), (float)(k.k2reg * pi4)};
}
__device__ __forceinline__ void pair_pb32(float dx, float dy, float dz, float qx, float qy, float qz, float qs,
                                          const PBConst32 &k, float &f0a, float &f0b, float &fx, float &fy,
                                          float &fz) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    const bool reg = !(r2 >= k.vr2);  // also r2 == 0
    const float ir = reg ? 0.0f : __builtin_amdgcn_rsqf(r2);
    const float r = r2 * ir;
    const float ir2 = ir * ir, ir3 = ir2 * ir;
    const float t = k.kappa * r;
    const float e = __builtin_amdgcn_exp2f(k.ck * r);
    const float et1 = fmaf(e, t, e);
    const float m = et1 * ir3;
    const float a3 = ir3 - m;
    const float g1 = fmaf(k.eps, m, -ir3);
    const float g4 = fmaf(k.inv_eps, m, -ir3);
    const float k2 = reg ? k.k2reg4pi : fmaf(-e, ir, ir);  // (1 - e) / r in one FMA
    const float b3 = (fmaf(et1, t + 2.0f, e) - 3.0f) * ir3 * ir2;
    const float dq = dx * qx + dy * qy + dz * qz;
    f0a = fmaf(g1, dq, f0a);
    f0b = fmaf(k2, qs, f0b);
    const float s = fmaf(b3, dq, g4 * qs);
    fx = fmaf(dx, s, fmaf(a3, qx, fx));
    fy = fmaf(dy, s, fmaf(a3, qy, fy));
    fz = fmaf(dz, s, fmaf(a3, qz, fz));
}

// FP64 near-field form of the same arithmetic (1/(4 pi) out of the sums,
// shared 1/r^3 product, fused accumulation), for the near-field work units.
// The regularisation test stays r < vacant radius on the square-rooted
// distance, as in pair_accum_pb (electric.cpp:58-101).
__device__ __forceinline__ void pair_pb64(double dx, double dy, double dz, double qx, double qy, double qz,
                                          double qs, const PBConst &k, double k2reg4pi, double &f0a, double &f0b,
                                          double &fx, double &fy, double &fz) {
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    const bool reg = r < k.vr;
    const double ir = reg ? 0.0 : recip(r);  // the discarded value may be inf or NaN (r = 0)
    const double ir2 = ir * ir, ir3 = ir2 * ir;
    const double t = k.kappa * r;
    const double e = exp(-t);
    const double et1 = fma(e, t, e);
    const double m = et1 * ir3;
    const double a3 = ir3 - m;
    const double g1 = fma(k.eps, m, -ir3);
    const double g4 = fma(k.inv_eps, m, -ir3);
    const double k2 = reg ? k2reg4pi : (1.0 - e) * ir;
    const double b3 = (fma(et1, t + 2.0, e) - 3.0) * ir3 * ir2;
    const double dq = dx * qx + dy * qy + dz * qz;
    f0a = fma(g1, dq, f0a);
    f0b = fma(k2, qs, f0b);
    const double s = fma(b3, dq, g4 * qs);
    fx = fma(dx, s, fma(a3, qx, fx));
    fy = fma(dy, s, fma(a3, qy, fy));
    fz = fma(dz, s, fma(a3, qz, fz));
}

template <bool POLY, typename T>
__device__ __forceinline__ void pair_accum_m(T dx, T dy, T dz, T qx, T qy, T qz, T qp, const PBConst &k, T &f0,
                                             T &fx, T &fy, T &fz) {
    if constexpr (POLY) pair_accum_poly<T>(dx, dy, dz, qx, qy, qz, qp, f0, fx, fy, fz);
    else pair_accum_pb<T>(dx, dy, dz, qx, qy, qz, qp, k, f0, fx, fy, fz);
}

// S_p(u, xbar_m) for m < P at one standardized coordinate u (bbfmm.h:408-419)
template <int P>
__device__ __forceinline__ void sp_values(double u, const ChebTab &tab, double *out) {
    double T[P];
    T[0] = 1.0;
    if (P > 1) T[1] = u;
#pragma unroll
    for (int k = 2; k < P; ++k) T[k] = 2.0 * u * T[k - 1] - T[k - 2];
#pragma unroll
    for (int m = 0; m < P; ++m) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) acc += T[k] * tab.Tbar[m][k];
        out[m] = (2.0 * acc - 1.0) * tab.invp;
    }
}

// stage timers: HIP events recorded on the plan's stream around each stage
enum Stage {
    kStP2M = 0, kStM2M, kStM2L, kStL2L, kStP2P, kStMatvec, kStOrth, kStRhs, kStPolarize, kStM2LX, kStMatvecMixed,
    kStM2LMixed, kStP2PMixed,
    kStCount
};

// Algorithmic work of one recorded stage interval, counted as in the
// roofline model of the matvec: 60 flops per near-field pair (vector ALU),
// 32 r^2 flops per V or X interaction (matrix cores), and the point and
// expansion bytes the translations move.
struct StageWork {
    double valu = 0, mfma = 0, bytes = 0;
};

struct StageTimer {
    bool on = false;
    unsigned mask = 0;  // bit s: stage s is recorded
    std::vector<hipEvent_t> pool;
    struct Rec {
        int stage;
        hipEvent_t a, b;
        StageWork w;
    };
    std::vector<Rec> pending;
    double ms[16] = {0};
    int64_t count[16] = {0};
    StageWork work[16];
    hipEvent_t take() {
        if (pool.empty()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
};

// Device blocks for plan-owned memory: released blocks go to a process-wide
// cache (per device) that serves later requests of similar size without
// hipMalloc; freeing a GB-sized block costs about a millisecond, and every
// solve of a request stream builds and releases a plan.  A released block
// must be idle (its stream synchronised).
hipError_t block_alloc(void **p, size_t bytes);
void block_release(void *p);
// hipMalloc for device buffers outside the block cache (level-set grids,
// request vectors, target sets): when it fails, the current device's idle
// cached blocks are freed and the allocation is tried once more
hipError_t dev_malloc_raw(void **p, size_t bytes);
template <typename T>
hipError_t dev_malloc(T **p, size_t bytes) {
    return dev_malloc_raw((void **)p, bytes);
}

// Device memory pool of one plan: chunks from hipMalloc handed out in
// 256-byte aligned pieces.  The persistent pool frees its chunks with the
// plan; the scratch pool is reset to a mark at the end of each build step
// (every use is ordered on the plan's stream, so reuse needs no wait).  A
// plan then holds a few chunks instead of a few hundred allocations, whose
// hipMalloc/hipFree calls cost about 10 ms per plan.
struct DevPool {
    struct Chunk {
        char *p;
        size_t size;
    };
    std::vector<Chunk> chunks;
    size_t cur = 0, off = 0;  // next free byte: chunk cur, offset off
    size_t min_chunk;
    int64_t bytes = 0;        // device memory held
    explicit DevPool(size_t mc) : min_chunk(mc) {}
    hipError_t alloc(void **ptr, size_t n) {
        n = n ? (n + 255) & ~(size_t)255 : 256;
        for (; cur < chunks.size(); ++cur, off = 0)
            if (off + n <= chunks[cur].size) {
                *ptr = chunks[cur].p + off;
                off += n;
                return hipSuccess;
            }
        const size_t sz = n > min_chunk ? n : min_chunk;
        char *p = nullptr;
        hipError_t e = block_alloc((void **)&p, sz);
        if (e != hipSuccess) return e;
        chunks.push_back({p, sz});
        bytes += (int64_t)sz;
        cur = chunks.size() - 1;
        off = n;
        *ptr = p;
        return hipSuccess;
    }
    std::pair<size_t, size_t> mark() const { return {cur, off}; }
    void release(std::pair<size_t, size_t> m) {
        cur = m.first;
        off = m.second;
    }
    void free_all() {
        for (const Chunk &c : chunks) block_release(c.p);
        chunks.clear();
        cur = off = 0;
        bytes = 0;
    }
};

// V-list interactions grouped by M2L operator class (level, |offset|)
struct VPairs {
    int npairs = 0;
    int *A = nullptr, *B = nullptr;     // source and target box of each pair, sorted by class
    unsigned char *mask = nullptr;      // reflection bits of the offset
    std::vector<int> cls_off;           // host: pairs of key c = 8 class + mask in [cls_off[c], cls_off[c+1])
    int *cls_off_d = nullptr;
    int nblocks = 0;                    // GEMM blocks of 64 pairs
    int *blk_cls = nullptr, *blk_j0 = nullptr;
    int hnblocks = 0;                   // FP16 GEMM blocks (256 V / 128 X pairs for p <= 5, 64 for p = 6,
                                        // 256 for p >= 7)
    int *hblk_cls = nullptr, *hblk_j0 = nullptr;
    int *hcls_off_d = nullptr;          // X list: pairs of class c in [hcls_off_d[c], hcls_off_d[c+1]); the FP16
                                        // blocks of the X list span whole classes (all 8 masks)
};

// a set of targets sorted into the leaves of a plan's tree
struct TargetSet {
    int n = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;  // sorted target coords
    const int *tbeg = nullptr, *tend = nullptr;             // per node ranges
    const int *leaves = nullptr;                            // leaves with targets
    int nleaves = 0;
    const int *m2l_work = nullptr;                          // nodes needing a local expansion
    int n_m2l_work = 0;
    const unsigned char *need_local = nullptr;              // per node
    const VPairs *vpairs = nullptr;                         // V interactions into those nodes
    bool xgemm = false;                                     // X list through the plan's class GEMMs
    const int *roff = nullptr;                              // per target leaf: merged near-field
    const int2 *ranges = nullptr;                           //   source ranges [roff[i], roff[i+1])
    const struct P2PWork *p2p = nullptr;                    // near-field work units over those ranges
};

// Near-field work units of a target set (at most 64 targets x kP2PChunk
// sources each, heaviest first) and the partial fields of split leaves
struct P2PWork {
    int nunits = 0;
    int4 *ua = nullptr;     // {target leaf index, first target, targets, partial slot or -1}
    int4 *ub = nullptr;     // {first range, offset in it, sources, first lane in the partial block}
    int2 *lsplit = nullptr; // per target leaf: {first partial slot, source chunks}
    double *part = nullptr; // 256 doubles (4 fields x 64 targets) per partial slot
    size_t npart = 0;
    std::vector<int64_t> pairs_prefix;  // host: target x source pairs of units [0, u), u = 0..nunits
                                        // (the plan's own points only; empty for target sets)
};

// ---------------------------------------------------------------------------
// Fused matvecs over a group of plans (concurrent solves of a server).  One
// launch per stage covers every member: its blocks are the members' blocks
// one after another, and each block finds its member in the launch's
// offset table, then reads that member's buffers from its MemberView (in
// device memory).  The members' own data are used in place; nothing is
// copied.  The kernels are the single-plan kernels' bodies, so a member's
// results do not depend on the group it runs in.
constexpr int kMaxGroup = 16;

struct MemberView {
    // sorted points, charges and fields
    const double *sx, *sy, *sz, *snx, *sny, *snz, *sw;
    const int *perm;
    double *Q, *F;
    int n;
    // tree
    int nnodes, n_leaves_ne, n_m2l_work;
    const int *leaves_ne, *sbeg, *send, *level, *flags, *first_child, *parent, *slot, *m2l_work;
    const double *ctr, *lrad;
    const unsigned char *need_local;
    double *M, *L;
    PBConst kc;
    double c1, c2;  // the combination's divisors (electric.cpp:382-383)
    // near field
    const int4 *ua, *ub;
    const int2 *ranges, *lsplit;  // lsplit null when no leaf is split
    double *part;
    // V and X lists, FP64 / FP32 class GEMMs
    const int *vblk, *vblk_j0, *vkey_off, *vA, *vB;
    const int *xblk, *xblk_j0, *xkey_off, *xA, *xB;
    const void *vops, *xops;      // FP64 or FP32 class operators
    const int *gtab, *stab;
    int MRP, KRP;
    // FP16 stage (p <= 5)
    const int *hvblk, *hvblk_j0, *keyslot, *hxblk, *hxblk_j0, *hxkey_off, *xiota, *xkeyslot;
    const unsigned char *xmask;
    const _Float16 *ops16;
    const int *esig;
    _Float16 *Mh, *Bx;
    int *ebox, *ebx;
    unsigned *Yv;
    int y_rows, vnpairs, xnpairs;
    const int *ytptr, *ytlist;
};

// this call's input and output vector of each member (kernel argument)
struct GroupXY {
    const double *x[kMaxGroup];
    double *y[kMaxGroup];
};

// one fused launch: members' block ranges [off[m], off[m+1]) and an
// optional per-member integer (a level's first node)
struct GroupLaunch {
    const MemberView *mv;
    const int *off;
    const int *aux;
    int k;
};

// Wave-uniform copies (scalar registers) of values a group kernel reads at
// block-uniform addresses of its member's tables.  Through pointers read
// from memory (the member's view) such loads land in vector registers, and
// loops bounded by them compile as divergent; readfirstlane makes them
// uniform again.  gu<G>(x) applies it in the group form of a kernel only, so
// the single-plan form compiles as before.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int2 uni(int2 x) { return make_int2(uni(x.x), uni(x.y)); }
__device__ __forceinline__ int4 uni(int4 x) { return make_int4(uni(x.x), uni(x.y), uni(x.z), uni(x.w)); }
__device__ __forceinline__ double uni(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)uni((int)b), hi = (unsigned)uni((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ __forceinline__ T *uni(T *p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = (unsigned)uni((int)u), hi = (unsigned)uni((int)(u >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}
template <bool G, typename T>
__device__ __forceinline__ T gu(T x) {
    if constexpr (G) return uni(x);
    else return x;
}

// the member of block b (off is nondecreasing, off[0] = 0), and b's index
// among that member's blocks (both wave-uniform)
__device__ __forceinline__ int group_member(const GroupLaunch &g, int b, int &local) {
    int m = 0;
    while (m + 1 < g.k && uni(g.off[m + 1]) <= b) ++m;
    local = uni(b - g.off[m]);
    return uni(m);
}

}  // namespace ibim

namespace ibim {
struct Batcher;
}

struct ibim_plan {
    ibim_params prm{};
    int n = 0;
    int p = 0, r = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    hipStream_t own_stream = nullptr;    // the plan's own stream (stream may be a caller's instead)
    hipEvent_t ev_switch = nullptr;      // orders work across a change of stream
    // FP16 mixed matvec: the near field runs on stream2 beside the M2L
    // reduction and L2L (forked after the M2L GEMMs, joined before L2P)
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool defer_reduce = false;           // the FP16 M2L stage leaves the reduction to the caller
    bool trace = false;                  // GMRES prints its iteration table to stdout
    ibim::Batcher *batcher = nullptr;    // a server's matvec batcher, while the plan's solve uses it
    ibim_plan *inner = nullptr;          // the order-prm.inner_order plan of the mixed inner solves (same points)

    // sorted geometry
    double *sx = nullptr, *sy = nullptr, *sz = nullptr;
    double *snx = nullptr, *sny = nullptr, *snz = nullptr, *sw = nullptr;
    int *perm = nullptr;

    // tree
    int nnodes = 0, nlevels = 0, depth = 0;
    double root_c[3] = {0, 0, 0}, root_r[3] = {0, 0, 0};
    std::vector<int> lvl_off;        // host, nlevels+1
    std::vector<double> lvl_rad;     // host, 3*nlevels
    double *d_lvl_rad = nullptr;
    double *ctr = nullptr;           // 3*nnodes
    uint64_t *nkey = nullptr;        // prefix key per node
    int *level = nullptr, *parent = nullptr, *first_child = nullptr, *slot = nullptr, *flags = nullptr;
    int *sbeg = nullptr, *send = nullptr;
    int *loff[4] = {nullptr, nullptr, nullptr, nullptr};   // U V W X offsets
    int *litem[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t ltot[4] = {0, 0, 0, 0};

    // work lists for the source set as targets
    int *leaves_ne = nullptr;   // non-empty leaves
    int n_leaves_ne = 0;
    int *m2l_work = nullptr;    // non-empty non-root nodes
    int n_m2l_work = 0;
    unsigned char *need_local = nullptr;

    // FMM buffers
    double *M = nullptr, *L = nullptr;   // nnodes * 4 * r
    double *Q = nullptr;                 // 4 * n sorted charges
    double *F = nullptr;                 // 4 * n sorted fields

    // GMRES workspace (allocated on first use)
    int gm_restart = 0;
    double *gm_V = nullptr;              // (restart+1) * 2n: the Krylov basis (FP64, inner solves too)
    double *gm_hpart = nullptr;          // (restart+1) * partials of the Gram-Schmidt projections
    double *gm_coef = nullptr;           // restart+1 projections of one Gram-Schmidt pass
    double *gm_w = nullptr, *gm_r = nullptr, *gm_x = nullptr, *gm_b = nullptr;
    double *gm_part = nullptr;           // partial sums
    double *gm_rr = nullptr, *gm_d = nullptr;  // refinement residual and correction
    double *gm_small = nullptr;          // H, s, cs, sn, scalars on device
    double *gm_host = nullptr;           // pinned host copies of those scalars (hipHostMalloc)

    ibim::StageTimer timer;

    // V-list M2L operators: n_cls classes of MRP x KRP (row major), FP64
    int n_cls = 0, MRP = 0, KRP = 0;
    double *ops = nullptr;
    float *ops32 = nullptr;              // FP32 copy for the mixed-precision matvec
    // FP16 M2L (IBIM_PREC_MIXED_F16): operators with one power-of-two scale
    // per row, multipoles with one per box, FP32 accumulation buffer
    _Float16 *ops16 = nullptr;           // one unfolded operator per (class, mask) key with pairs, V keys
                                         // (slot keyslot[key]) then X keys (slot xkeyslot[key])
    int *esig = nullptr;                 // per (operator, row) exponent
    int *keyslot = nullptr, *xkeyslot = nullptr;
    bool x16 = false;                    // the X list runs in the FP16 stage
    bool f16_ready = false;              // every FP16 stage buffer above is built
    _Float16 *Mh = nullptr;              // nnodes * (4r rounded up to 64), per box scaled
    int *ebox = nullptr;                 // per box exponent
    unsigned *Yv = nullptr;              // V then X interaction results as FP24 (3 dwords per 4 rows),
                                         // pairs * (4r rounded up to 32) values
    int y_rows = 0;                      // pair rows of Yv; one spare row follows them
    size_t yv_dwords = 0;                // dwords allocated for Yv
    _Float16 *Bx = nullptr;              // X pairs: reflected FP16 source multipoles, one row per pair
    int *ebx = nullptr, *xiota = nullptr; // X pairs: the source box exponent; 0, 1, 2, ...
    int *ytptr = nullptr, *ytlist = nullptr; // rows of Yv of target node b: ytlist[ytptr[b] .. ytptr[b+1])
    int *gtab = nullptr, *stab = nullptr; // gather / scatter tables of the 8 reflections
    ibim::VPairs vp;                     // V pairs for the points as their own targets
    // X list as GEMMs: pairs grouped by class (level of the target, level
    // difference, offset), one operator per class (key = 8 class, no reflection)
    ibim::VPairs xp;
    int n_xcls = 0;
    double *xops = nullptr;
    float *xops32 = nullptr;
    const int *p2p_roff = nullptr;       // merged near-field ranges of the non-empty leaves
    const int2 *p2p_ranges = nullptr;
    ibim::P2PWork p2p;                   // near-field work units of the plan's own points

    ibim::ChebTab tab{};
    ibim::PBConst kc{};

    std::vector<void *> allocs;          // large allocations to free
    ibim::DevPool pool{(size_t)32 << 20};  // the plan's other buffers
    ibim::DevPool spool{(size_t)32 << 20}; // build-time scratch
    int64_t footprint = 0;
    ibim_tree_stats stats{};
};

namespace ibim {
// allocation helper for buffers that live as long as the plan: large ones
// get their own allocation, the rest come from the plan's pool; counts the
// device memory held
template <typename T>
int dalloc(ibim_plan *pl, T **ptr, size_t count) {
    *ptr = nullptr;
    if (count == 0) count = 1;
    const size_t bytes = count * sizeof(T);
    hipError_t e;
    if (bytes >= ((size_t)16 << 20)) {
        e = block_alloc((void **)ptr, bytes);
        if (e == hipSuccess) {
            pl->allocs.push_back((void *)*ptr);
            pl->footprint += (int64_t)bytes;
        }
    } else {
        const int64_t held = pl->pool.bytes;
        e = pl->pool.alloc((void **)ptr, bytes);
        pl->footprint += pl->pool.bytes - held;
    }
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
        return IBIM_ERR_NOMEM;
    }
    return IBIM_OK;
}

// build-time scratch of `bytes` bytes from the plan's scratch pool, valid
// until the enclosing ScratchScope ends
template <typename T>
hipError_t smalloc(ibim_plan *pl, T **ptr, size_t bytes) {
    return pl->spool.alloc((void **)ptr, bytes);
}
struct ScratchScope {
    ibim_plan *pl;
    std::pair<size_t, size_t> m;
    explicit ScratchScope(ibim_plan *p) : pl(p), m(p->spool.mark()) {}
    ~ScratchScope() { pl->spool.release(m); }
};

// stages implemented in the .hip files
int build_tree(ibim_plan *pl, const double *d_xyz, const double *d_nrm, const double *d_w);
int build_chebyshev(ibim_plan *pl);
int fmm_forward(ibim_plan *pl, const TargetSet &ts, double *F, int nt, bool mixed = false);
int form_charges(ibim_plan *pl, const double *d_x);
int combine(ibim_plan *pl, const double *d_x, double *d_y);
// Per-call device temporaries of a target-set evaluation: device memory from
// dev_malloc (freed by release()), or consecutive 256-byte aligned pieces of
// a caller-provided arena, in which case nothing is allocated and a request
// past the arena's end fails
struct TmpAlloc {
    char *arena = nullptr;
    size_t cap = 0, used = 0;
    std::vector<void *> owned;
    int alloc(void **p, size_t bytes);
    template <typename T>
    int get(T **p, size_t count) {
        return alloc((void **)p, (count ? count : 1) * sizeof(T));
    }
    void release();
};
int make_target_set(ibim_plan *pl, int m, const double *d_txyz, TargetSet &ts, TmpAlloc &tmp, int **perm_out);
TargetSet source_targets(ibim_plan *pl);
int rhs_device(ibim_plan *pl, int m, const double *d_axyz, const double *d_aq, double *d_b);
int gmres_device(ibim_plan *pl, const double *d_b, double *d_x, int restart, int maxit, double tol,
                 ibim_solve_info *info);
int compute_stats(ibim_plan *pl);
int build_m2l_operators(ibim_plan *pl);
int build_vpairs(ibim_plan *pl, const unsigned char *need, VPairs &vp, std::vector<void *> *tmp_owner);
int m2l_gemm(ibim_plan *pl, const VPairs &vp, bool mixed);
int m2l_reduce(ibim_plan *pl);  // the FP16 stage's per-target reduction (after defer_reduce)
int build_xpairs(ibim_plan *pl);
int m2l_gemm_x(ibim_plan *pl, bool mixed);
int build_m2l_operators_f32(ibim_plan *pl);
int build_m2l_operators_f16(ibim_plan *pl);
int build_p2p_ranges(ibim_plan *pl, TargetSet &ts, P2PWork &work, std::vector<void *> *tmp_owner);

// scoped stage timing on the plan's stream (no-op when profiling is off),
// or on a fused group's timer and stream
struct StageScope {
    StageTimer *tm;
    int stage;
    hipEvent_t a = nullptr;
    hipStream_t st;
    StageWork w;
    StageScope(ibim_plan *p, int s, hipStream_t on = nullptr) : StageScope(p->timer, s, on ? on : p->stream) {}
    StageScope(StageTimer &t, int s, hipStream_t on) : tm(&t), stage(s), st(on) {
        if (tm->on && ((tm->mask >> stage) & 1u)) {
            a = tm->take();
            if (a) (void)hipEventRecord(a, st);
        }
    }
    // the algorithmic work done inside this interval
    void add(double valu, double mfma, double bytes) {
        w.valu += valu;
        w.mfma += mfma;
        w.bytes += bytes;
    }
    ~StageScope() {
        if (a) {
            hipEvent_t b = tm->take();
            if (b) {
                (void)hipEventRecord(b, st);
                tm->pending.push_back({stage, a, b, w});
            } else {
                tm->pool.push_back(a);
            }
        }
    }
};
// resolve a timer's recorded intervals into its per-stage sums (waits for
// their end events)
int timer_collect(StageTimer &t);
// collect, then copy the per-stage sums into *out (if out), and zero them (if
// reset)
int timer_report(StageTimer &t, ibim_stage_report *out, int reset);

// algorithmic work of near-field work units [u0, u1) of the plan's own
// points (60 flops per pair), and of one whole matvec of the plan
double near_field_flops(const ibim_plan *pl, int u0, int u1);
StageWork matvec_work(const ibim_plan *pl, bool mixed);

// A fused matvec over k plans of one device, Chebyshev order and kernel
// family, all in the same matvec form (FP64, FP32-mixed, or FP16-mixed with
// p <= 5).  Its launch tables (block offsets of every member in every fused
// launch) are built once, for the plans as they are; the group must be
// destroyed before any member is.
struct MatvecGroup {
    struct Launch {
        int off = 0, aux = -1;  // entry of the member offsets (k + 1) and per-member values (k) in tab
        int blocks = 0;         // blocks (tiles for the GEMMs) over all members
    };
    int k = 0, p = 0, device = 0, form = IBIM_PREC_FP64;
    bool poly = false;
    std::vector<ibim_plan *> pl;
    std::vector<int> tab;                  // host copy of the launch tables
    MemberView *d_mv = nullptr;            // device: the members' views
    int *d_tab = nullptr;                  // device: tab
    Launch charges, p2m, zero_l, near_early, near_rest, near_all, m_to_half, x_reflect, v_gemm, x_gemm, v_reduce,
        l2p, combine;
    std::vector<Launch> m2m, l2l;          // one per level step (M2M from each member's finest level up)
    StageWork work;                        // algorithmic work of one fused matvec (sum over members)
    double near_early_flops = 0, near_rest_flops = 0, m2l_flops = 0;
    hipStream_t st = nullptr, st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_done = nullptr;
    std::vector<hipEvent_t> ev_in;         // per member: its input is ready on its stream
    StageTimer timer;                      // stage timers of the fused launches
};
// the members' streams -> group -> the members' streams: a fused matvec
// ordered after what each member's stream holds, and before what it gets next
int group_matvec_ordered(MatvecGroup *g, const double *const *x, double *const *y);

// one plan's matvec on its own stream, and the matvec a solve asks for: the
// batcher's (fused with other solves' matvecs) while the plan has one
int matvec_single(ibim_plan *pl, const double *d_x, double *d_y, bool mixed);
int batch_matvec(Batcher *b, ibim_plan *pl, const double *d_x, double *d_y, bool mixed);
// checks the plans and builds the group on its own streams; a plan that
// cannot join (form, order, device, missing tables) gives IBIM_ERR_ARG
int group_create(ibim_plan *const *plans, int k, bool mixed, MatvecGroup **out);
void group_destroy(MatvecGroup *g);
// y_m = A_m x_m for every member, on the group's streams: the caller orders
// the members' inputs before it (group stream waits) and their later work
// after g->ev_done
int group_matvec(MatvecGroup *g, const double *const *x, double *const *y);
// the fused M2L stage (V and X lists) of a group, on g->st; FP16 form: the
// per-target reduction is left to group_m2l_reduce
int group_m2l(MatvecGroup *g);
int group_m2l_reduce(MatvecGroup *g);
}  // namespace ibim
