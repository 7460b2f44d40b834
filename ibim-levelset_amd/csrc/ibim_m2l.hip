// ibim_m2l.hip -- the V-list M2L as batched dense GEMMs on the matrix cores.
//
// Reference: downPass, V list (bbfmm.h:547-552): L_B += K(grid_B, grid_A) M_A
// with the kernel matrix evaluated entry by entry for every pair of
// same-level boxes.  Two same-level boxes at offset o = (c_A - c_B)/(2 r_l)
// (o in [-3,3]^3, |o|_inf >= 2: 316 offsets) always see the same kernel
// matrix, so the 16 fused kernels become one (4r x 4r) operator per level and
// offset, applied to every pair with that offset as one GEMM
//     Y = K_{l,o} [M_A1 M_A2 ...],   L_Bj += Y_j.
// Reflections fold the 316 offsets onto the 56 with o >= 0: with sigma the
// sign of each component of o, the operator of o is the operator of |o| with
// the Chebyshev node index reversed along each reflected axis and the
// x/y/z field components negated along it (K1, K4 odd and K2, K3 even in d),
// so the gather of M_A and the scatter into L_B apply that permutation and
// those signs instead of storing 8x more operators.
//
// Operators are built once per plan, in FP64.  The GEMM runs on FP64 MFMA
// (v_mfma_f64_16x16x4f64): 64 x 64 output tiles, 4 waves of 32 x 32, the
// operator rows and the gathered multipoles staged through LDS; results are
// added into L with hardware FP64 atomics (pairs with the same target box
// come from different offsets, i.e. different GEMMs).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "ibim_fp24.h"
#include "ibim_internal.h"

namespace ibim {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

inline int nblk(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

// canonical offset index of |o| in [0,3]^3 with max >= 2 (56 of them)
__host__ __device__ inline int canon_index(int ax, int ay, int az) {
    int idx = 0;
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                if (x < 2 && y < 2 && z < 2) continue;
                if (x == ax && y == ay && z == az) return idx;
                ++idx;
            }
    return -1;
}
constexpr int kNCanon = 56;

__host__ __device__ inline void canon_offset(int idx, int &ax, int &ay, int &az) {
    int c = 0;
    for (int z = 0; z < 4; ++z)
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) {
                if (x < 2 && y < 2 && z < 2) continue;
                if (c == idx) {
                    ax = x;
                    ay = y;
                    az = z;
                    return;
                }
                ++c;
            }
}

// V pairs of every target box that needs a local expansion: count / fill
template <int PASS>
__global__ void k_vpairs(int nnodes, const unsigned char *__restrict__ need, const int *__restrict__ voff,
                         const int *__restrict__ vit, const int *__restrict__ flags, const int *__restrict__ level,
                         const double *__restrict__ ctr, const double *__restrict__ lrad, int *__restrict__ cnt,
                         const int *__restrict__ off, int *__restrict__ pA, int *__restrict__ pB,
                         unsigned char *__restrict__ pm, int *__restrict__ pkey) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nnodes) return;
    int c = 0;
    if (b != 0 && need[b]) {
        const int lev = level[b];
        const double *rd = lrad + 3 * lev;
        for (int q = voff[b]; q < voff[b + 1]; ++q) {
            const int a = vit[q];
            if (flags[a] & kEmpty) continue;
            if (PASS == 1) {
                int o[3];
                for (int d = 0; d < 3; ++d) o[d] = (int)rint((ctr[3 * a + d] - ctr[3 * b + d]) / (2.0 * rd[d]));
                const int m = (o[0] < 0 ? 1 : 0) | (o[1] < 0 ? 2 : 0) | (o[2] < 0 ? 4 : 0);
                const int ci = canon_index(abs(o[0]), abs(o[1]), abs(o[2]));
                const int j = off[b] + c;
                pA[j] = a;
                pB[j] = b;
                pm[j] = (unsigned char)m;
                pkey[j] = ((lev - 2) * kNCanon + ci) * 8 + m;  // grouped by class, then reflection
            }
            ++c;
        }
    }
    if (PASS == 0) cnt[b] = c;
}

__global__ void k_permute_pairs(int np, const int *__restrict__ order, const int *__restrict__ a_in,
                                const int *__restrict__ b_in, const unsigned char *__restrict__ m_in,
                                int *__restrict__ a_out, int *__restrict__ b_out, unsigned char *__restrict__ m_out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= np) return;
    const int s = order[j];
    a_out[j] = a_in[s];
    b_out[j] = b_in[s];
    m_out[j] = m_in[s];
}

// operator entries of class (level, |o|): row g*r + t, column f*r + s
template <int P>
__global__ void k_build_ops(int lev0, int ncls, const double *__restrict__ lrad, ChebTab tab, PBConst kc, int MRP,
                            int KRP, double *__restrict__ ops) {
    constexpr int R = P * P * P;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)R * R;
    if (e >= per * ncls) return;
    const int c = (int)(e / per);
    const int t = (int)((e % per) / R), s = (int)(e % R);
    const int lev = lev0 + c / kNCanon;
    int o[3];
    canon_offset(c % kNCanon, o[0], o[1], o[2]);
    const double *rd = lrad + 3 * lev;
    const int ti[3] = {t % P, (t / P) % P, t / (P * P)};
    const int si[3] = {s % P, (s / P) % P, s / (P * P)};
    double d[3];
    for (int q = 0; q < 3; ++q) {
        double xs = 0.0, xt = 0.0;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            if (i == si[q]) xs = tab.xbar[i];
            if (i == ti[q]) xt = tab.xbar[i];
        }
        d[q] = rd[q] * (2.0 * o[q] + xs - xt);
    }
    // column f = input field, row g = output field; pair_accum with unit
    // charges gives each column of the 4 x 4 block
    double blk[4][4];
    for (int f = 0; f < 4; ++f) {
        double q[4] = {0, 0, 0, 0};
        q[f] = 1.0;
        double f0 = 0, fx = 0, fy = 0, fz = 0;
        pair_accum<double>(d[0], d[1], d[2], q[0], q[1], q[2], q[3], kc, f0, fx, fy, fz);
        blk[0][f] = f0;
        blk[1][f] = fx;
        blk[2][f] = fy;
        blk[3][f] = fz;
    }
    double *op = ops + (size_t)c * MRP * KRP;
    for (int g = 0; g < 4; ++g)
        for (int f = 0; f < 4; ++f) op[(size_t)(g * R + t) * KRP + f * R + s] = blk[g][f];
}

template <int P>
__device__ __forceinline__ int flip_node(int s, int m) {
    int i = s % P, j = (s / P) % P, k = s / (P * P);
    if (m & 1) i = P - 1 - i;
    if (m & 2) j = P - 1 - j;
    if (m & 4) k = P - 1 - k;
    return (k * P + j) * P + i;
}

// one (class, 64 pairs, 64 rows) tile per block; T = double (FP64 MFMA) or
// float (FP32 MFMA, FP32 operators, multipoles rounded to FP32 on load)
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // result register q of lane l holds row (l >> 4) + 4q
    static __device__ __forceinline__ int row(int l, int q) { return (l >> 4) + 4 * q; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // result register q of lane l holds row 4 (l >> 4) + q
    static __device__ __forceinline__ int row(int l, int q) { return 4 * (l >> 4) + q; }
};

// Gather table of one reflection mask m: column k = f*r + s of the operator
// reads multipole element f*r + flip_m(s) of the source box, negated for the
// reflected x/y/z fields; the scatter uses the same permutation on the rows
// with the output-field signs.  gtab[m][k] = index | (negate << 31).
template <int P>
__global__ void k_gather_table(int *__restrict__ gtab, int *__restrict__ stab) {
    constexpr int R = P * P * P;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 8 * 4 * R) return;
    const int m = e / (4 * R), k = e % (4 * R);
    const int f = k / R, s = k % R;
    const int idx = f * R + flip_node<P>(s, m);
    const bool neg_in = f < 3 && ((m >> f) & 1);
    const bool neg_out = f > 0 && ((m >> (f - 1)) & 1);
    gtab[e] = idx | (neg_in ? (int)0x80000000 : 0);
    stab[e] = idx | (neg_out ? (int)0x80000000 : 0);
}

// One (class, mask, 64 pairs, 64 operator rows) tile per block, 4 waves of
// 32 x 32.  Computed as Y^T = X^T K^T so that a result register holds 16
// consecutive operator rows of one pair: the atomic adds into L_B of one
// instruction hit 16 neighbouring elements of one box.  The next K-chunk of
// the operator and of the gathered multipoles is loaded into registers while
// the current one is multiplied from LDS.
// GROUP (XL: the X list): the tiles of every member of a fused group, count
// = their total, the XCD renumbering over the whole launch.
template <int P, typename T, bool GROUP = false, bool XL = false>
// four waves per SIMD: the register budget (at most 128) keeps the chunk
// prefetch without spilling
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_m2l_gemm(const int *__restrict__ blk_key, const int *__restrict__ blk_j0,
                                                  const int *__restrict__ key_off, const int *__restrict__ pA,
                                                  const int *__restrict__ pB, const int *__restrict__ gtab,
                                                  const int *__restrict__ stab, const T *__restrict__ ops, int MRP,
                                                  int KRP, const double *__restrict__ M, double *__restrict__ L,
                                                  int first, int count, int per_xcd, GroupLaunch g = {}) {
    constexpr int R = P * P * P;
    constexpr int R4 = 4 * R;
    constexpr int KC = 32;
    typedef typename Mfma<T>::v4 v4;
    __shared__ T sA[64][KC + 1];   // operator rows x k
    __shared__ T sB[64][KC + 1];   // pairs x k
    __shared__ int sPA[64], sPB[64];
    __shared__ int sS[64];
    // Workgroups go to the 8 XCDs round robin; renumber so that each XCD
    // gets one contiguous range of tiles.  Tiles are ordered (class, mask,
    // pair tile, row tile): the row tiles of a pair tile, and the pair tiles
    // of one operator class, then run on one XCD and share its L2.  A launch
    // covers tiles first .. first + count - 1.
    const int local = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (local >= count) return;
    int logical = first + local;
    if constexpr (GROUP) {
        const MemberView &v = g.mv[group_member(g, local, logical)];
        blk_key = uni(XL ? v.xblk : v.vblk);
        blk_j0 = uni(XL ? v.xblk_j0 : v.vblk_j0);
        key_off = uni(XL ? v.xkey_off : v.vkey_off);
        pA = uni(XL ? v.xA : v.vA);
        pB = uni(XL ? v.xB : v.vB);
        gtab = uni(v.gtab);
        stab = uni(v.stab);
        ops = uni((const T *)(XL ? v.xops : v.vops));
        MRP = uni(v.MRP);
        KRP = uni(v.KRP);
        M = uni(v.M);
        L = uni(v.L);
    }
    const int nrt = MRP / 64;
    const int tile = logical / nrt;
    const int tid = threadIdx.x;
    const int key = gu<GROUP>(blk_key[tile]);
    const int cls = key >> 3, m = key & 7;
    const int j0 = gu<GROUP>(blk_j0[tile]);
    const int jend = gu<GROUP>(key_off[key + 1]);
    const int row0 = (logical % nrt) * 64;
    if (tid < 64) {
        const int j = j0 + tid;
        const bool ok = j < jend;
        sPA[tid] = ok ? pA[j] : -1;
        sPB[tid] = ok ? pB[j] : -1;
        const int row = row0 + tid;
        sS[tid] = row < R4 ? stab[m * R4 + row] : 0;
    }
    const int w = tid >> 6, l = tid & 63;
    const int wr = (w & 1) * 32;  // operator rows of this wave
    const int wc = (w >> 1) * 32; // pairs of this wave
    v4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{T(0), T(0), T(0), T(0)};
    const T *Aop = ops + (size_t)cls * MRP * KRP + (size_t)row0 * KRP;
    const int *gt = gtab + m * R4;
    // each thread moves 8 elements of each tile per chunk: k = kc + (tid & 31)
    // of rows (tid >> 5) + 8 i.  Loads are unconditional (clamped indices,
    // masked values) so that all 16 of them are in flight together.
    const int kk = tid & (KC - 1), r8 = tid >> 5;
    T ra[8], rb[8];
    auto load_chunk = [&](int kc) {
        const int k = kc + kk;
        const int g = gt[min(k, R4 - 1)];
        const int gi = g & 0x7fffffff;
#pragma unroll
        for (int i = 0; i < 8; ++i) ra[i] = Aop[(size_t)(r8 + 8 * i) * KRP + k];
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = M[(size_t)max(sPA[r8 + 8 * i], 0) * R4 + gi];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const bool ok = sPA[r8 + 8 * i] >= 0 && k < R4;
            rb[i] = ok ? (T)(g < 0 ? -v[i] : v[i]) : T(0);
        }
    };
    __syncthreads();  // sPA, sPB, sS
    load_chunk(0);
    for (int kc = 0; kc < KRP; kc += KC) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            sA[r8 + 8 * i][kk] = ra[i];
            sB[r8 + 8 * i][kk] = rb[i];
        }
        __syncthreads();
        if (kc + KC < KRP) load_chunk(kc + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            T xf[2], kf[2];
#pragma unroll
            for (int pi = 0; pi < 2; ++pi) xf[pi] = sB[wc + pi * 16 + (l & 15)][kk + (l >> 4)];
#pragma unroll
            for (int ri = 0; ri < 2; ++ri) kf[ri] = sA[wr + ri * 16 + (l & 15)][kk + (l >> 4)];
#pragma unroll
            for (int pi = 0; pi < 2; ++pi)
#pragma unroll
                for (int ri = 0; ri < 2; ++ri) acc[pi][ri] = Mfma<T>::op(xf[pi], kf[ri], acc[pi][ri]);
        }
        __syncthreads();
    }
    // result of tile (pi, ri): register q of lane l holds pair Mfma::row(l, q)
    // and operator row l & 15
#pragma unroll
    for (int pi = 0; pi < 2; ++pi)
#pragma unroll
        for (int ri = 0; ri < 2; ++ri) {
            const int lrow = wr + ri * 16 + (l & 15);
            const int row = row0 + lrow;
            if (row >= R4) continue;
            const int st = sS[lrow];
            const int dst = st & 0x7fffffff;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = wc + pi * 16 + Mfma<T>::row(l, q);
                const int b = sPB[col];
                if (b < 0) continue;
                double v = (double)acc[pi][ri][q];
                if (st < 0) v = -v;
                unsafeAtomicAdd(&L[(size_t)b * R4 + dst], v);
            }
        }
}

__global__ void k_to_float(const double *__restrict__ a, int64_t n, float *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

template <int P>
int build_ops_p(ibim_plan *pl) {
    constexpr int R = P * P * P;
    const int ncls = pl->n_cls;
    if (ncls == 0) return IBIM_OK;
    const int64_t total = (int64_t)R * R * ncls;
    int rc;
    if ((rc = dalloc(pl, &pl->gtab, 8 * 4 * R)) || (rc = dalloc(pl, &pl->stab, 8 * 4 * R))) return rc;
    k_gather_table<P><<<nblk(8 * 4 * R), 256, 0, pl->stream>>>(pl->gtab, pl->stab);
    IBIM_HIP_CHECK(hipMemsetAsync(pl->ops, 0, (size_t)ncls * pl->MRP * pl->KRP * sizeof(double), pl->stream));
    k_build_ops<P><<<nblk(total), 256, 0, pl->stream>>>(2, ncls, pl->d_lvl_rad, pl->tab, pl->kc, pl->MRP, pl->KRP,
                                                        pl->ops);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

// ------------------------------------------------- FP16 V-list GEMM ---
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Pair results are kept as FP24: the upper 24 bits of the FP32 value,
// rounded to nearest (relative error below 2^-16, far under the FP16
// products' own), four consecutive rows packed into three dwords.
__device__ __forceinline__ unsigned fp24(float x) { return (__float_as_uint(x) + 0x80u) >> 8; }
__device__ __forceinline__ void store4_fp24(unsigned *p, float a, float b, float c, float d) {
    const unsigned A = fp24(a), B = fp24(b), C = fp24(c), D = fp24(d);
    p[0] = A | (B << 24);
    p[1] = (B >> 8) | (C << 16);
    p[2] = (C >> 16) | (D << 8);
}
__device__ __forceinline__ f4v load4_fp24(const unsigned *p) {
    const unsigned x = p[0], y = p[1], z = p[2];
    return f4v{__uint_as_float(x << 8), __uint_as_float(((x >> 24) | (y << 8)) << 8),
               __uint_as_float(((y >> 16) | (z << 16)) << 8), __uint_as_float(z & 0xffffff00u)};
}

// e such that every |v| of a set with maximum vmax satisfies |v| 2^-e < 1
// (0 for an all-zero set)
__device__ __forceinline__ int scale_exp(double vmax) {
    if (!(vmax > 0.0)) return 0;
    int e;
    (void)frexp(vmax, &e);
    return e;
}

// per box: FP16 multipoles with one power-of-two scale (every matvec)
// rows of Mh are KP = R4 rounded up to 64 long, zero beyond R4
template <bool GROUP = false>
__global__ __launch_bounds__(64) void k_m_to_half(const double *__restrict__ M, int R4, int KP,
                                                  _Float16 *__restrict__ Mh, int *__restrict__ ebox,
                                                  GroupLaunch g = {}) {
    size_t a = blockIdx.x;
    if constexpr (GROUP) {  // (every box of every member)
        int bid;
        const MemberView &v = g.mv[group_member(g, blockIdx.x, bid)];
        a = bid;
        M = uni(v.M);
        Mh = uni(v.Mh);
        ebox = uni(v.ebox);
    }
    const double *src = M + a * R4;
    double mx = 0.0;
    for (int k = threadIdx.x; k < R4; k += 64) mx = fmax(mx, fabs(src[k]));
    for (int o = 32; o; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    const int e = scale_exp(mx);
    if (threadIdx.x == 0) ebox[a] = e;
    for (int k = threadIdx.x; k < KP; k += 64) Mh[a * KP + k] = k < R4 ? (_Float16)ldexp(src[k], -e) : (_Float16)0;
}

// Unfolded FP16 operators: slot s holds the key (class c, reflection mask m)
// and U[i][j] = so * si * K_c[s(i)][g(j)] 2^-esig[i], where g and s are the
// mask's gather and scatter permutations (reflections, their own inverses)
// and so, si the output / input signs, so that L_b += U M_a needs neither
// permutation.  One block of 64 per (class, class row R): the row is read
// once into LDS and written, permuted and signed, as row i = s(R) of every
// slot of the class (the row's scale exponent does not depend on the mask).
// cls_slots[8 q + m] is the slot of (cls_list[q], m), or -1.
__global__ __launch_bounds__(64) void k_unfold_half(const int *__restrict__ cls_list,
                                                    const int *__restrict__ cls_slots,
                                                    const double *__restrict__ ops, int MRP, int KRP, int R4,
                                                    const int *__restrict__ gtab, const int *__restrict__ stab,
                                                    _Float16 *__restrict__ ops16u, int *__restrict__ esig) {
    extern __shared__ double srow[];  // KRP doubles
    const int R = blockIdx.x, q = blockIdx.y;
    const int c = cls_list[q];
    const int *sl8 = cls_slots + 8 * q;
    if (R >= R4) {  // padding row: zero in every slot
        for (int m = 0; m < 8; ++m) {
            const int sl = sl8[m];
            if (sl < 0) continue;
            _Float16 *dst = ops16u + ((size_t)sl * MRP + R) * KRP;
            for (int j = threadIdx.x; j < KRP; j += 64) dst[j] = (_Float16)0;
            if (threadIdx.x == 0) esig[(size_t)sl * MRP + R] = 0;
        }
        return;
    }
    const double *src = ops + ((size_t)c * MRP + R) * KRP;
    double mx = 0.0;
    for (int k = threadIdx.x; k < R4; k += 64) {
        const double v = src[k];
        srow[k] = v;
        mx = fmax(mx, fabs(v));
    }
    for (int o = 32; o; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    const int e = scale_exp(mx);
    __syncthreads();
    for (int m = 0; m < 8; ++m) {
        const int sl = sl8[m];
        if (sl < 0) continue;
        const int sr = stab[m * R4 + R];
        const int i = sr & 0x7fffffff;
        const bool nout = sr < 0;
        _Float16 *dst = ops16u + ((size_t)sl * MRP + i) * KRP;
        if (threadIdx.x == 0) esig[(size_t)sl * MRP + i] = e;
        for (int j = threadIdx.x; j < KRP; j += 64) {
            _Float16 v = (_Float16)0;
            if (j < R4) {
                const int k = gtab[m * R4 + j] & 0x7fffffff;
                const bool nin = gtab[m * R4 + k] < 0;
                const double x = ldexp(srow[k], -e);
                v = (_Float16)(nout != nin ? -x : x);
            }
            dst[j] = v;
        }
    }
}

typedef __attribute__((address_space(3))) void lds_void;

// one direct global-to-LDS load: lane i's 16 B at p go to LDS byte address
// base + 16 i (base wave-uniform).  Issued as inline assembly, so that the
// compiler's wait insertion does not treat later LDS reads as dependent on
// it: the caller orders the data with wait_vm and a barrier.
__device__ __forceinline__ void lds_load16(const void *p, lds_void *base) {
    const unsigned m = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)base);
    asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(p), "s"(m) : "memory", "m0");
}

// s_waitcnt for at most N outstanding vector memory operations of this wave
// (expcnt and lgkmcnt not waited for)
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// The FP16 V-list GEMM (v_mfma_f32_16x16x32_f16, FP32 accumulation) on
// unfolded operators (one per V-list (class, mask) key with pairs,
// k_unfold_half; for the X list one per class, the reflection applied to
// the multipoles and the results instead): a block takes 4 PPW pairs of one
// key, wave w the pairs [PPW w, PPW w + PPW), PPW = 16 (register budget at
// p = 6; p <= 5 runs k_m2l_gemm_h32).  Each wave
// loads its pairs' MFMA B-fragments (8 consecutive k of one pair = 16 B of
// Mh) for all k straight into registers once.  The operator streams through
// an LDS ring of NST chunks of 64 rows x KC k, filled by direct global-to-LDS
// loads NST - 1 chunks ahead of use, so the HBM latency of the operator
// (which all blocks of a key request at about the same time) is hidden; one
// pass over the operator serves the block's pairs.  A chunk row's 16-byte
// segments are stored XOR-swizzled by the row index, which keeps the MFMA
// A-fragment reads free of bank conflicts.  A lane's four accumulators are
// four consecutive rows of one pair: results, scaled by 2^(esig + ebox), are
// stored as FP24, 12 B per lane, as Y[pair][row], rows padded to YS = R4
// rounded up to 32 (whole 128 B lines per 64-row tile); k_v_reduce sums every
// target's pairs into L.
template <int P, int PPW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_m2l_gemm_hu(
    const int *__restrict__ blk_key, const int *__restrict__ blk_j0, const int *__restrict__ key_off,
    const int *__restrict__ pA, const int *__restrict__ keyslot, const _Float16 *__restrict__ ops16u,
    const int *__restrict__ esig, int MRP, const _Float16 *__restrict__ Mh, const int *__restrict__ ebox,
    unsigned *__restrict__ Y, int nblocks, int per_xcd) {
    constexpr int R = P * P * P;
    constexpr int R4 = 4 * R;
    constexpr int KP = (R4 + 63) / 64 * 64;
    constexpr int KS = KP / 32;                  // MFMA k-steps
    constexpr int KC = KP % 128 == 0 ? 128 : 64; // k per operator chunk
    constexpr int NKC = KP / KC;
    constexpr int SEGS = KC / 8;                 // 16-byte segments per chunk row
    constexpr int RPI = 64 / SEGS;               // chunk rows per LDS-load instruction (64 lanes x 16 B)
    constexpr int DPW = SEGS / 4;                // LDS-load instructions per wave per chunk
    constexpr int NST = 4;                       // chunks in the LDS ring
    constexpr int CH = 64 * KC;                  // halves per chunk
    constexpr int PS = PPW / 16;                 // 16-pair subtiles of a wave
    constexpr int NPB = 4 * PPW;                 // pairs per block
    constexpr int YS = (R4 + 31) / 32 * 32;      // row stride of Y
    static_assert(KP % KC == 0 && DPW >= 1 && R4 % 4 == 0, "tile shape");
    constexpr int MR = (R4 + 63) / 64 * 64;      // operator rows (padded)
    __shared__ __attribute__((aligned(1024))) _Float16 sA[NST][CH];
    __shared__ int sEA[NPB];
    __shared__ __attribute__((aligned(16))) int sE[MR];  // the operator's row exponents
    const int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (tile >= nblocks) return;
    const int tid = threadIdx.x;
    const int key = blk_key[tile];
    const int j0 = blk_j0[tile];
    const int jend = key_off[key + 1];
    const int sl = keyslot[key];
    const int w = tid >> 6, l = tid & 63;
    const _Float16 *Aop = ops16u + (size_t)sl * MRP * KP;
    const int nrt = MRP / 64;
    const int NC = nrt * NKC;
    // chunk c into ring slot c % NST: instruction i of wave w fills rows
    // RPI (DPW w + i) .. + RPI - 1; lane l's 16 B land at row r = RPI (DPW w
    // + i) + l / SEGS, position l % SEGS, and hold segment (l % SEGS) ^ (r %
    // SEGS) of that row
    auto issue = [&](int c) {
        const int rt = c / NKC, kc = c - rt * NKC;
        _Float16 *dst = sA[c % NST];
#pragma unroll
        for (int i = 0; i < DPW; ++i) {
            const int r = RPI * (DPW * w + i) + l / SEGS;
            const int sg = (l % SEGS) ^ (r % SEGS);
            lds_load16(Aop + (size_t)(rt * 64 + r) * KP + kc * KC + 8 * sg, (lds_void *)(dst + (DPW * w + i) * 512));
        }
    };
#pragma unroll
    for (int c = 0; c < NST - 1; ++c)
        if (c < NC) issue(c);
    if (tid < NPB) {
        const int j = j0 + tid;
        sEA[tid] = j < jend ? ebox[pA[j]] : 0;
    }
    // (global loads inside the chunk loop would make the compiler wait for
    // every outstanding load, the LDS ring's included)
    for (int i = tid; i < MR; i += 256) sE[i] = esig[(size_t)sl * MRP + i];
    // B-fragments of this wave's pairs, all k
    h8 xr[KS][PS];
#pragma unroll
    for (int pi = 0; pi < PS; ++pi) {
        const int j = j0 + PPW * w + 16 * pi + (l & 15);
        const bool ok = j < jend;
        const _Float16 *src = Mh + (size_t)(ok ? pA[j] : 0) * KP + 8 * (l >> 4);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) xr[ks][pi] = ok ? *(const h8 *)(src + 32 * ks) : h8{};
    }
    int c = 0;
    for (int rt = 0; rt < nrt; ++rt) {
        f4v acc[PS][4];
#pragma unroll
        for (int a = 0; a < PS; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc, ++c) {
            // this wave's loads of chunk c are done once at most the loads of
            // chunks c + 1 .. c + NST - 2 are outstanding (a lower bound on
            // what was issued after them: stores may add more); the barrier
            // then publishes every wave's part and retires the reads of chunk
            // c - 1, whose ring slot the next issue refills
            if (c + NST - 2 < NC) wait_vm<DPW * (NST - 2)>();
            else wait_vm<0>();
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            if (c + NST - 1 < NC) issue(c + NST - 1);
            const _Float16 *sa = sA[c % NST];
#pragma unroll
            for (int ks = 0; ks < KC / 32; ++ks) {
                h8 kf[4];
                const int sg = (4 * ks + (l >> 4)) ^ ((l & 15) % SEGS);
#pragma unroll
                for (int ri = 0; ri < 4; ++ri) kf[ri] = *(const h8 *)(sa + (16 * ri + (l & 15)) * KC + 8 * sg);
#pragma unroll
                for (int pi = 0; pi < PS; ++pi)
#pragma unroll
                    for (int ri = 0; ri < 4; ++ri)
                        acc[pi][ri] = __builtin_amdgcn_mfma_f32_16x16x32_f16(kf[ri], xr[kc * (KC / 32) + ks][pi],
                                                                             acc[pi][ri], 0, 0, 0);
            }
        }
        // register q of lane l holds row 4 (l >> 4) + q of the row tile, pair
        // l & 15 of the subtile; R4 is a multiple of 4, so a lane's four rows
        // are all below R4 or all above
#pragma unroll
        for (int ri = 0; ri < 4; ++ri) {
            const int row = rt * 64 + 16 * ri + 4 * (l >> 4);
            if (row >= R4) continue;
            const int4 er = *(const int4 *)&sE[row];
            const int e0 = er.x, e1 = er.y, e2 = er.z, e3 = er.w;
#pragma unroll
            for (int pi = 0; pi < PS; ++pi) {
                const int col = PPW * w + 16 * pi + (l & 15);
                if (j0 + col >= jend) continue;
                const int ea = sEA[col];
                const f4v a = acc[pi][ri];
                store4_fp24(Y + fp24_offset(j0 + col, YS, row), ldexpf(a[0], e0 + ea), ldexpf(a[1], e1 + ea),
                            ldexpf(a[2], e2 + ea), ldexpf(a[3], e3 + ea));
            }
        }
    }
}

// The same GEMM on v_mfma_f32_32x32x16_f16, which on gfx950 does twice the
// work of the 16x16x32 form in the same issue time: a wave holds the
// B-fragments of 32 pairs (one 32-column tile) for all k and multiplies them
// with the operator chunk as two 32-row A-tiles per 16-k step.  Operands: A
// lane l holds rows l & 31, k 8 (l >> 5) .. + 7 of the step; B lane l holds
// pair l & 31, the same k; accumulator j of lane l is row 8 (j >> 2) + 4 (l
// >> 5) + (j & 3), pair l & 31.  For p <= 5 (the B-fragments of 32 pairs over
// all k fit the register budget).
// waves per block of the V- and X-list launches: a block reads its key's
// operator once for 32 NW pairs, and a V key has about a thousand pairs
// while an X key has a few dozen
constexpr int kH32WavesV = 8, kH32WavesX = 4;

// (GROUP: the pair blocks of every member of a fused group, nblocks their
// total, the XCD renumbering over the whole launch; 8-wave blocks run the V
// list, 4-wave blocks the X list, whose rows follow the V rows in Yv)
template <int P, int NW, bool GROUP = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW >= 8 ? 2 : 1))) void k_m2l_gemm_h32(
    const int *__restrict__ blk_key, const int *__restrict__ blk_j0, const int *__restrict__ key_off,
    const int *__restrict__ pA, const int *__restrict__ keyslot, const _Float16 *__restrict__ ops16u,
    const int *__restrict__ esig, int MRP, const _Float16 *__restrict__ Mh, const int *__restrict__ ebox,
    unsigned *__restrict__ Y, int yspare, int nblocks, int per_xcd, GroupLaunch g = {}) {
    constexpr int R = P * P * P;
    constexpr int R4 = 4 * R;
    constexpr int KP = (R4 + 63) / 64 * 64;
    constexpr int KS = KP / 16;                  // MFMA k-steps
    // k per operator chunk: 128 where a row tile has at least DIST chunks of
    // it (p = 5), else 64 (p = 4: four chunks per row tile, so its stores
    // take the counted, LDS-staged path below)
    constexpr int KC = KP % 128 == 0 && KP / 128 >= 3 ? 128 : 64;
    constexpr int NKC = KP / KC;
    constexpr int SEGS = KC / 8;                 // 16-byte segments per chunk row
    constexpr int RPI = 64 / SEGS;               // chunk rows per LDS-load instruction (64 lanes x 16 B)
    constexpr int DPW = SEGS / NW;               // LDS-load instructions per wave per chunk
    constexpr int NST = 5;                       // chunks in the LDS ring
    constexpr int DIST = NST - 2;                // chunks loading ahead of the one being read
    constexpr int KSC = KC / 16;                 // MFMA k-steps per chunk
    constexpr int PD = 3;                        // A-fragment prefetch depth (k-steps)
    constexpr int CH = 64 * KC;                  // halves per chunk
    constexpr int NPB = 32 * NW;                 // pairs per block (32 per wave)
    constexpr int YS = (R4 + 31) / 32 * 32;      // row stride of Y
    constexpr int MR = (R4 + 63) / 64 * 64;      // operator rows (padded)
    static_assert(KP % KC == 0 && DPW >= 1 && DPW * NW == SEGS && R4 % 4 == 0 && P <= 5, "tile shape");
    static_assert(KSC > PD && (NKC * KSC) % (PD + 1) == 0, "pipeline shape");
    // Every row tile's results leave as exactly NSTO stores per wave (the
    // padded rows and the lanes without a pair write too, the latter to the
    // spare row), so the loads of a chunk can be waited for exactly while
    // the stores are in flight: vmcnt completes in issue order.  Otherwise
    // the waits also drain the stores and the chunks loading behind them.
    // COUNTED tiles pass through a per-wave LDS stage so that each pair's
    // 64 rows (48 dwords of FP24) leave as 16-byte stores of consecutive
    // lanes instead of 12-byte pieces 32 pairs apart.
    constexpr bool COUNTED = YS >= MR && NKC >= DIST;
    constexpr bool STAGED = COUNTED && NW == 8;   // (the LDS of 4-wave blocks stays at two blocks per CU)
    constexpr int PS = 52;                        // staged dwords per pair (48 used; 16-byte rows)
    constexpr int NSTO = STAGED ? 6 : 8;          // 32 pairs x 12 16-byte pieces / 64 lanes, or 8 x 12 bytes
    __shared__ __attribute__((aligned(16))) unsigned sStg[STAGED ? NW * 32 * PS : 1];
    __shared__ __attribute__((aligned(1024))) _Float16 sA[NST][CH];
    __shared__ int sEA[NPB];
    __shared__ __attribute__((aligned(16))) int sE[MR];  // the operator's row exponents
    int tile = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (tile >= nblocks) return;
    if constexpr (GROUP) {
        static_assert(kH32WavesV != kH32WavesX, "the block size tells the lists apart");
        constexpr bool XL = NW == kH32WavesX;
        constexpr int YS1 = (4 * P * P * P + 31) / 32 * 32;
        const MemberView &v = g.mv[group_member(g, tile, tile)];
        blk_key = uni(XL ? v.hxblk : v.hvblk);
        blk_j0 = uni(XL ? v.hxblk_j0 : v.hvblk_j0);
        key_off = uni(XL ? v.hxkey_off : v.vkey_off);
        pA = uni(XL ? v.xiota : v.vA);
        keyslot = uni(XL ? v.xkeyslot : v.keyslot);
        ops16u = uni(v.ops16);
        esig = uni(v.esig);
        MRP = uni(v.MRP);
        Mh = uni(XL ? v.Bx : v.Mh);
        ebox = uni(XL ? v.ebx : v.ebox);
        const int npv = uni(v.vnpairs), rows = uni(v.y_rows);
        Y = uni(v.Yv) + (XL ? fp24_offset(npv, YS1, 0) : 0);
        yspare = XL ? rows - npv : rows;
    }
    const int tid = threadIdx.x;
    const int key = gu<GROUP>(blk_key[tile]);
    const int j0 = gu<GROUP>(blk_j0[tile]);
    const int jend = gu<GROUP>(key_off[key + 1]);
    const int sl = gu<GROUP>(keyslot[key]);
    const int w = tid >> 6, l = tid & 63;
    const _Float16 *Aop = ops16u + (size_t)sl * MRP * KP;
    const int nrt = MRP / 64;
    const int NC = nrt * NKC;
    // chunk c into ring slot c % NST, rows XOR-swizzled as in k_m2l_gemm_hu
    auto issue = [&](int c) {
        const int rt = c / NKC, kc = c - rt * NKC;
        _Float16 *dst = sA[c % NST];
#pragma unroll
        for (int i = 0; i < DPW; ++i) {
            const int r = RPI * (DPW * w + i) + l / SEGS;
            const int sg = (l % SEGS) ^ (r % SEGS);
            lds_load16(Aop + (size_t)(rt * 64 + r) * KP + kc * KC + 8 * sg, (lds_void *)(dst + (DPW * w + i) * 512));
        }
    };
#pragma unroll
    for (int c = 0; c < DIST; ++c)
        if (c < NC) issue(c);
    if (tid < NPB) {
        const int j = j0 + tid;
        sEA[tid] = j < jend ? ebox[pA[j]] : 0;
    }
    for (int i = tid; i < MR; i += 64 * NW) sE[i] = esig[(size_t)sl * MRP + i];
    // B-fragments of this wave's 32 pairs, all k
    h8 xr[KS];
    {
        const int j = j0 + 32 * w + (l & 31);
        const bool ok = j < jend;
        const _Float16 *src = Mh + (size_t)(ok ? pA[j] : 0) * KP + 8 * (l >> 5);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) xr[ks] = ok ? *(const h8 *)(src + 16 * ks) : h8{};
    }
    // Chunk cc becomes readable at arrive(cc): each wave waits for its own
    // part of the chunk, then the block barrier; the chunk DIST ahead is then
    // issued into the slot of chunk cc - 2, which every wave has finished.
    // arrive(cc + 1) runs PD k-steps before the end of chunk cc, so the
    // A-fragments of the next chunk are read while this one multiplies: the
    // fragments of k-step g + PD are requested at step g, across chunk and
    // row-tile boundaries.  (Stores only add to vmcnt and make a wait
    // stricter, never looser.)
    auto arrive = [&](int cc, bool after_stores) {
        if (cc + DIST - 1 >= NC) wait_vm<0>();
        else if (COUNTED && after_stores) wait_vm<DPW * (DIST - 1) + NSTO>();
        else wait_vm<DPW * (DIST - 1)>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        if (cc + DIST < NC) issue(cc + DIST);
    };
    auto frag = [&](int cc, int ks, h8 *k) {
        const _Float16 *sa = sA[cc % NST];
        const int sg = (2 * ks + (l >> 5)) ^ ((l & 31) % SEGS);
        k[0] = *(const h8 *)(sa + (l & 31) * KC + 8 * sg);
        k[1] = *(const h8 *)(sa + (32 + (l & 31)) * KC + 8 * sg);
    };
    h8 F[PD + 1][2];  // fragments of k-steps g .. g + PD (g mod (PD + 1))
    arrive(0, false);
#pragma unroll
    for (int q = 0; q < PD; ++q) frag(0, q, F[q]);
    for (int rt = 0; rt < nrt; ++rt) {
        f16v acc[2];
        acc[0] = f16v{};
        acc[1] = f16v{};
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            const int c = rt * NKC + kc;
#pragma unroll
            for (int ks = 0; ks < KSC; ++ks) {
                const int gi = kc * KSC + ks;  // step within the row tile (a multiple of PD + 1 long)
                // the previous row tile's stores are younger than the loads
                // of chunk c + 1 when kc < DIST
                if (ks == KSC - PD && c + 1 < NC) arrive(c + 1, rt > 0 && kc < DIST);
                // (past the last chunk the read repeats a resident slot and
                // is not used: no branch, so the LDS wait counts stay exact)
                const int c2 = min(c + (ks + PD) / KSC, NC - 1);
                frag(c2, (ks + PD) % KSC, F[(gi + PD) % (PD + 1)]);
                // (sched_barrier: the scheduler would sink the reads next to
                // their use to save registers)
                __builtin_amdgcn_sched_barrier(0);
                const h8 b = xr[gi];
                acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(F[gi % (PD + 1)][0], b, acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(F[gi % (PD + 1)][1], b, acc[1], 0, 0, 0);
            }
        }
        // rows 8 q + 4 (l >> 5) .. + 3 of each 32-row half, pair l & 31; R4
        // is a multiple of 4, so a lane's four rows are all below R4 or all
        // above
        const int col = 32 * w + (l & 31);
        const bool has_pair = j0 + col < jend;
        if constexpr (STAGED) {
            const int ea = sEA[col];  // 0 without a pair
            unsigned *stg = sStg + w * 32 * PS;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r4 = 8 * h + 2 * q + (l >> 5);  // 4-row group within the tile
                    const int4 er = *(const int4 *)&sE[rt * 64 + 4 * r4];
                    const f16v &a = acc[h];
                    store4_fp24(stg + (l & 31) * PS + 3 * r4, ldexpf(a[4 * q], er.x + ea),
                                ldexpf(a[4 * q + 1], er.y + ea), ldexpf(a[4 * q + 2], er.z + ea),
                                ldexpf(a[4 * q + 3], er.w + ea));
                }
            // (a wave's LDS operations execute in order: no barrier; the
            // pieces go two at a time to bound the live registers)
#pragma unroll
            for (int i = 0; i < NSTO; i += 2) {
                uint4 v[2];
                size_t o[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int t = 64 * (i + u) + l, pp = t / 12, k = t - 12 * pp;
                    v[u] = *(const uint4 *)(stg + pp * PS + 4 * k);
                    const int jj = j0 + 32 * w + pp;
                    o[u] = fp24_offset(jj < jend ? jj : yspare, YS, 64 * rt) + 4 * k;
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) *(uint4 *)(Y + o[u]) = v[u];
                __builtin_amdgcn_sched_barrier(0);
            }
        } else if (COUNTED) {
            const int ea = sEA[col];  // 0 without a pair
            unsigned *yp = Y + fp24_offset(has_pair ? j0 + col : yspare, YS, 0);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = rt * 64 + 32 * h + 8 * q + 4 * (l >> 5);  // < MR <= YS
                    const int4 er = *(const int4 *)&sE[row];
                    const f16v &a = acc[h];
                    store4_fp24(yp + row / 4 * 3, ldexpf(a[4 * q], er.x + ea), ldexpf(a[4 * q + 1], er.y + ea),
                                ldexpf(a[4 * q + 2], er.z + ea), ldexpf(a[4 * q + 3], er.w + ea));
                }
        } else if (has_pair) {
            const int ea = sEA[col];
            unsigned *yp = Y + fp24_offset(j0 + col, YS, 0);
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = rt * 64 + 32 * h + 8 * q + 4 * (l >> 5);
                    if (row >= R4) continue;
                    const int4 er = *(const int4 *)&sE[row];
                    const f16v &a = acc[h];
                    store4_fp24(yp + row / 4 * 3, ldexpf(a[4 * q], er.x + ea), ldexpf(a[4 * q + 1], er.y + ea),
                                ldexpf(a[4 * q + 2], er.z + ea), ldexpf(a[4 * q + 3], er.w + ea));
                }
        }
    }
}

// The FP16 V-list GEMM for p >= 7, where a pair's results (4 p^3 >= 1372
// rows) are too many to keep per pair as the p <= 6 stage does (FP24 rows of
// 61.5M pairs at p = 8 would take 378 GB).  Same operands: the unfolded
// operators of the (class, mask) keys with their row exponents, and the
// per-box scaled FP16 multipoles.  A block computes kHtRows operator rows x
// kHtPairs pairs of one key over all k; both operands stream through an LDS
// ring of NST chunks of KC k, filled by direct global-to-LDS loads NST - 1
// chunks ahead (the pairs' multipole rows gathered by source box), segments
// XOR-swizzled by row as in k_m2l_gemm_hu.  Eight waves of 64 rows x 64
// pairs.  The MFMA takes the pairs as its first operand and the operator
// rows as its second, so result register j of lane l holds operator row
// l & 31 of pair 8 (j >> 2) + 4 (l >> 5) + (j & 3): the results, scaled by
// 2^(esig + ebox), are added into L in FP64 by atomics, each half-wave
// covering 32 consecutive elements of one box.  Tiles are ordered (key,
// pair tile, row tile), so the row tiles of a pair tile run side by side and
// share its multipole rows in L2.  REFLECT (the X list): keys are classes,
// the operators unreflected, the multipole rows those of k_x_reflect_half
// (one per pair), and result row r goes to element s(r) of the target with
// the sign of stab[mask * R4 + r] (the reflection is its own inverse and
// keeps each field's block).
constexpr int kHtRows = 128, kHtPairs = 256;

template <int P, bool REFLECT>
__global__ __launch_bounds__(512) void k_m2l_gemm_ht(
    const int *__restrict__ blk_key, const int *__restrict__ blk_j0, const int *__restrict__ key_off,
    const int *__restrict__ pA, const int *__restrict__ pB, const int *__restrict__ keyslot,
    const _Float16 *__restrict__ ops16u, const int *__restrict__ esig, int MRP, const _Float16 *__restrict__ Mh,
    const int *__restrict__ ebox, const unsigned char *__restrict__ pmask, const int *__restrict__ stab,
    double *__restrict__ L, int first, int count, int per_xcd) {
    constexpr int R = P * P * P;
    constexpr int R4 = 4 * R;
    constexpr int KP = (R4 + 63) / 64 * 64;   // operator columns and multipole row length
    constexpr int KC = 64;                    // k per chunk
    constexpr int NKC = KP / KC;
    constexpr int SEGS = KC / 8;              // 16-byte segments per chunk row
    constexpr int RPI = 64 / SEGS;            // chunk rows per LDS-load instruction
    constexpr int DA = kHtRows / RPI / 8;     // operator LDS-load instructions per wave per chunk
    constexpr int DB = kHtPairs / RPI / 8;    // multipole LDS-load instructions per wave per chunk
    constexpr int NST = 3;                    // chunks in the ring
    static_assert(KP % KC == 0 && DA >= 1 && DB >= 1 && R4 % 4 == 0 && NKC >= NST, "tile shape");
    __shared__ __attribute__((aligned(1024))) _Float16 sA[NST][kHtRows * KC];
    __shared__ __attribute__((aligned(1024))) _Float16 sB[NST][kHtPairs * KC];
    __shared__ int sSrc[kHtPairs], sDst[kHtPairs], sEb[kHtPairs];
    __shared__ int sMask[REFLECT ? kHtPairs : 1];
    __shared__ int sE[kHtRows];
    const int local = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (local >= count) return;
    const int logical = first + local;
    const int nrt = MRP / kHtRows;
    const int tile = logical / nrt, rt = logical - tile * nrt;
    const int key = blk_key[tile];
    const int j0 = blk_j0[tile];
    const int jend = key_off[key + 1];
    const int sl = keyslot[key];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    if (tid < kHtPairs) {
        const int j = j0 + tid;
        const bool ok = j < jend;
        const int a = ok ? pA[j] : 0;  // (a pair-less lane reads box 0's row; its results are dropped)
        sSrc[tid] = a;
        sDst[tid] = ok ? pB[j] : -1;
        sEb[tid] = ok ? ebox[a] : 0;
        if constexpr (REFLECT) sMask[tid] = ok ? pmask[j] : 0;
    } else if (tid < kHtPairs + kHtRows) {
        sE[tid - kHtPairs] = esig[(size_t)sl * MRP + rt * kHtRows + (tid - kHtPairs)];
    }
    __syncthreads();
    const _Float16 *Aop = ops16u + ((size_t)sl * MRP + (size_t)rt * kHtRows) * KP;
    // chunk c into ring slot c % NST: instruction i of wave w fills rows
    // RPI (D w + i) .. + RPI - 1 of the slot; lane l's 16 B land at row
    // r = RPI (D w + i) + l / SEGS, position l % SEGS, and hold segment
    // (l % SEGS) ^ (r % SEGS) of that row
    auto issue = [&](int c) {
        _Float16 *da = sA[c % NST], *db = sB[c % NST];
#pragma unroll
        for (int i = 0; i < DA; ++i) {
            const int r = RPI * (DA * w + i) + l / SEGS;
            const int sg = (l % SEGS) ^ (r % SEGS);
            lds_load16(Aop + (size_t)r * KP + c * KC + 8 * sg, (lds_void *)(da + (DA * w + i) * 512));
        }
#pragma unroll
        for (int i = 0; i < DB; ++i) {
            const int q = RPI * (DB * w + i) + l / SEGS;
            const int sg = (l % SEGS) ^ (q % SEGS);
            lds_load16(Mh + (size_t)sSrc[q] * KP + c * KC + 8 * sg, (lds_void *)(db + (DB * w + i) * 512));
        }
    };
#pragma unroll
    for (int c = 0; c < NST - 1; ++c) issue(c);
    const int wr = (w & 1) * 64;   // this wave's operator rows in the tile
    const int wp = (w >> 1) * 64;  // this wave's pairs in the tile
    f16v acc[2][2];                // [pair group][row group]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f16v{};
    for (int c = 0; c < NKC; ++c) {
        // this wave's loads of chunk c are done once at most those of chunks
        // c + 1 .. c + NST - 2 are outstanding; the barrier then publishes
        // every wave's part and retires the reads of chunk c - 1, whose slot
        // the next issue refills
        if (c + NST - 2 < NKC) wait_vm<(DA + DB) * (NST - 2)>();
        else wait_vm<0>();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        if (c + NST - 1 < NKC) issue(c + NST - 1);
        const _Float16 *sa = sA[c % NST], *sb = sB[c % NST];
#pragma unroll
        for (int ks = 0; ks < KC / 16; ++ks) {
            const int seg = 2 * ks + (l >> 5);
            h8 fp[2], fr[2];
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                const int q = wp + 32 * g + (l & 31);
                fp[g] = *(const h8 *)(sb + q * KC + 8 * (seg ^ (q % SEGS)));
                const int r = wr + 32 * g + (l & 31);
                fr[g] = *(const h8 *)(sa + r * KC + 8 * (seg ^ (r % SEGS)));
            }
#pragma unroll
            for (int pg = 0; pg < 2; ++pg)
#pragma unroll
                for (int rg = 0; rg < 2; ++rg)
                    acc[pg][rg] = __builtin_amdgcn_mfma_f32_32x32x16_f16(fp[pg], fr[rg], acc[pg][rg], 0, 0, 0);
        }
    }
#pragma unroll
    for (int rg = 0; rg < 2; ++rg) {
        const int rl = wr + 32 * rg + (l & 31);  // row in the tile
        const int row = rt * kHtRows + rl;
        if (row >= R4) continue;                 // padding rows
        const int er = sE[rl];
#pragma unroll
        for (int pg = 0; pg < 2; ++pg)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int q = wp + 32 * pg + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
                const int b = sDst[q];
                if (b < 0) continue;
                const double v = ldexp((double)acc[pg][rg][j], er + sEb[q]);
                if constexpr (REFLECT) {
                    const int sr = stab[sMask[q] * R4 + row];
                    unsafeAtomicAdd(&L[(size_t)b * R4 + (sr & 0x7fffffff)], sr < 0 ? -v : v);
                } else {
                    unsafeAtomicAdd(&L[(size_t)b * R4 + row], v);
                }
            }
    }
}

// X pair j: its source multipole reflected by the pair's mask (the gather
// table's permutation and input signs), so that the class operator applies
// unreflected; and the source box's exponent
template <bool GROUP = false>
__global__ __launch_bounds__(64) void k_x_reflect_half(const int *__restrict__ A, const unsigned char *__restrict__ mask,
                                                       const int *__restrict__ gtab, int R4, int KP,
                                                       const _Float16 *__restrict__ Mh, const int *__restrict__ ebox,
                                                       _Float16 *__restrict__ Bx, int *__restrict__ ebx,
                                                       GroupLaunch g = {}) {
    int j = blockIdx.x;
    if constexpr (GROUP) {  // (every X pair of every member)
        const MemberView &v = g.mv[group_member(g, blockIdx.x, j)];
        A = uni(v.xA);
        mask = uni(v.xmask);
        gtab = uni(v.gtab);
        Mh = uni(v.Mh);
        ebox = uni(v.ebox);
        Bx = uni(v.Bx);
        ebx = uni(v.ebx);
    }
    const int a = gu<GROUP>(A[j]), m = gu<GROUP>((int)mask[j]);
    const _Float16 *src = Mh + (size_t)a * KP;
    _Float16 *dst = Bx + (size_t)j * KP;
    constexpr int U = 8;  // elements per lane per pass, their loads in flight together
    for (int k0 = 0; k0 < KP; k0 += 64 * U) {
        int g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + 64 * u + threadIdx.x;
            g[u] = k < R4 ? gtab[m * R4 + k] : -1;
        }
        _Float16 x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = g[u] == -1 ? (_Float16)0 : src[g[u] & 0x7fffffff];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + 64 * u + threadIdx.x;
            if (k < KP) dst[k] = g[u] != -1 && g[u] < 0 ? -x[u] : x[u];
        }
    }
    if (threadIdx.x == 0) ebx[j] = ebox[a];
}

// L_b += sum over the V pairs (a, b) of the pair's result: one block per
// target node, a thread per four rows (one pass: 4 blockDim >= R4), four
// pairs' rows in flight per step.  The rows of X pairs (row index >= npv,
// after the V rows of each target) hold the unreflected class operator's
// result: element e of the target takes row element stab[m][e] of the
// pair's mask m, with its sign; such a row is decoded into LDS first.
// Fixed order: the sum is deterministic.
template <int R4, bool GROUP = false>
__global__ __launch_bounds__(256) void k_v_reduce(const int *__restrict__ work, const int *__restrict__ tptr,
                                                  const int *__restrict__ tlist, const unsigned *__restrict__ Y,
                                                  int npv, const unsigned char *__restrict__ xmask,
                                                  const int *__restrict__ stab, double *__restrict__ L,
                                                  GroupLaunch g = {}) {
    constexpr int YS = (R4 + 31) / 32 * 32;
    static_assert(R4 % 4 == 0 && R4 / 4 <= 256, "one pass");
    __shared__ __attribute__((aligned(16))) float sx[R4];
    int bid = blockIdx.x;
    if constexpr (GROUP) {  // (the nodes needing a local expansion, every member)
        const MemberView &v = g.mv[group_member(g, blockIdx.x, bid)];
        work = uni(v.m2l_work);
        tptr = uni(v.ytptr);
        tlist = uni(v.ytlist);
        Y = uni(v.Yv);
        npv = uni(v.vnpairs);
        xmask = uni(v.xmask);
        stab = uni(v.stab);
        L = uni(v.L);
    }
    const int b = gu<GROUP>(work[bid]);
    const int q0 = gu<GROUP>(tptr[b]), q1 = gu<GROUP>(tptr[b + 1]);
    if (q0 == q1) return;
    int qx = q1;  // the X rows close the list
    while (qx > q0 && gu<GROUP>(tlist[qx - 1]) >= npv) --qx;
    const int e = 4 * threadIdx.x;
    const bool act = e < R4;
    f4v ax = {0.f, 0.f, 0.f, 0.f};
    // X rows, up to XB at a time: this thread's four elements of each are
    // loaded together, then passed through LDS for the permutation
    constexpr int XB = 4;
    for (int qb = qx; qb < q1; qb += XB) {
        f4v xv[XB];
        int xm[XB];
#pragma unroll
        for (int i = 0; i < XB; ++i) {
            const int q = min(qb + i, q1 - 1);
            const int row = gu<GROUP>(tlist[q]);
            xm[i] = gu<GROUP>((int)xmask[row - npv]);
            xv[i] = act ? load4_fp24(Y + fp24_offset(row, YS, e)) : f4v{};
        }
#pragma unroll
        for (int i = 0; i < XB; ++i) {
            if (qb + i >= q1) break;  // (block-uniform)
            __syncthreads();  // the previous row is read
            if (act) *(f4v *)&sx[e] = xv[i];
            __syncthreads();
            if (act) {
                const int *st = stab + xm[i] * R4 + e;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int sr = st[u];
                    const float v = sx[sr & 0x7fffffff];
                    ax[u] += sr < 0 ? -v : v;
                }
            }
        }
    }
    if (act) {
        f4v a0 = {0.f, 0.f, 0.f, 0.f}, a1 = a0, a2 = a0, a3 = a0;
        int q = q0;
        for (; q + 3 < qx; q += 4) {
            a0 += load4_fp24(Y + fp24_offset(gu<GROUP>(tlist[q]), YS, e));
            a1 += load4_fp24(Y + fp24_offset(gu<GROUP>(tlist[q + 1]), YS, e));
            a2 += load4_fp24(Y + fp24_offset(gu<GROUP>(tlist[q + 2]), YS, e));
            a3 += load4_fp24(Y + fp24_offset(gu<GROUP>(tlist[q + 3]), YS, e));
        }
        for (; q < qx; ++q) a0 += load4_fp24(Y + fp24_offset(gu<GROUP>(tlist[q]), YS, e));
        const f4v t = ((a0 + a1) + (a2 + a3)) + ax;
        double *dst = L + (size_t)b * R4 + e;
        dst[0] += (double)t[0];
        dst[1] += (double)t[1];
        dst[2] += (double)t[2];
        dst[3] += (double)t[3];
    }
}

// The blocked FP64/FP32 GEMM over pair tiles x row tiles.  A dispatch's
// grid is a 32-bit count of work-items, so one launch covers at most
// kGemmChunk tiles of 256 threads; at p = 8 a 58M-point tube has 31M tiles.
constexpr int64_t kGemmChunk = int64_t(1) << 23;

template <int P, typename T>
int launch_gemm(const VPairs &vp, const ibim_plan *pl, const T *ops, hipStream_t st) {
    const int64_t total = (int64_t)vp.nblocks * (pl->MRP / 64);
    if (total >= (int64_t(1) << 31)) {
        set_error("M2L: too many pair tiles for one plan");
        return IBIM_ERR_STATE;
    }
    for (int64_t first = 0; first < total; first += kGemmChunk) {
        const int count = (int)std::min(kGemmChunk, total - first);
        const int per_xcd = (count + 7) / 8;
        k_m2l_gemm<P, T><<<8 * per_xcd, 256, 0, st>>>(vp.blk_cls, vp.blk_j0, vp.cls_off_d, vp.A, vp.B, pl->gtab,
                                                      pl->stab, ops, pl->MRP, pl->KRP, pl->M, pl->L, (int)first, count,
                                                      per_xcd);
    }
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

template <int P>
int gemm_p(ibim_plan *pl, const VPairs &vp, bool mixed) {
    // (the FP16 stage also evaluates the X list: it runs when either list
    // has pairs)
    if (vp.nblocks == 0 && (&vp != &pl->vp || pl->xp.npairs == 0)) return IBIM_OK;
    constexpr int R4 = 4 * P * P * P;
    hipStream_t st = pl->stream;
    if constexpr (P <= 6) {
        if (mixed && pl->prm.precision == IBIM_PREC_MIXED_F16 && pl->f16_ready && &vp == &pl->vp) {
            constexpr int KP = (R4 + 63) / 64 * 64;
            constexpr int YS = (R4 + 31) / 32 * 32;
            // every store lands in a pair row or the spare row after them
            // (rows y_rows), values below YS: check that against the buffer
            if (fp24_offset(pl->y_rows, YS, YS - 4) + 3 > pl->yv_dwords) {
                set_error("FP16 M2L: pair-result rows exceed their buffer");
                return IBIM_ERR_STATE;
            }
            k_m_to_half<<<pl->nnodes, 64, 0, st>>>(pl->M, R4, KP, pl->Mh, pl->ebox);
            // V pairs, then (x16) the X pairs into the rows after them.
            // key_off: pair range of each block key; pA / Mh / ebox: the
            // pairs' multipole rows and exponents; Y: the launch's first
            // pair row; yspare: the spare row after all pair rows, relative
            // to Y
            auto gemm = [&](const VPairs &pp, const int *key_off, const int *pA, const _Float16 *Mh, const int *ebox,
                            const int *keyslot, unsigned *Y, int yspare, bool vlist) {
                const int pxcd = (pp.hnblocks + 7) / 8;
                if constexpr (P <= 5) {
                    if (vlist)
                        k_m2l_gemm_h32<P, kH32WavesV><<<8 * pxcd, 64 * kH32WavesV, 0, st>>>(
                            pp.hblk_cls, pp.hblk_j0, key_off, pA, keyslot, pl->ops16, pl->esig, pl->MRP, Mh, ebox, Y,
                            yspare, pp.hnblocks, pxcd);
                    else
                        k_m2l_gemm_h32<P, kH32WavesX><<<8 * pxcd, 64 * kH32WavesX, 0, st>>>(
                            pp.hblk_cls, pp.hblk_j0, key_off, pA, keyslot, pl->ops16, pl->esig, pl->MRP, Mh, ebox, Y,
                            yspare, pp.hnblocks, pxcd);
                } else
                    k_m2l_gemm_hu<P, 16><<<8 * pxcd, 256, 0, st>>>(pp.hblk_cls, pp.hblk_j0, key_off, pA, keyslot,
                                                                   pl->ops16, pl->esig, pl->MRP, Mh, ebox, Y,
                                                                   pp.hnblocks, pxcd);
            };
            if (vp.hnblocks > 0)
                gemm(vp, vp.cls_off_d, vp.A, pl->Mh, pl->ebox, pl->keyslot, pl->Yv, pl->y_rows, true);
            const VPairs &xp = pl->xp;
            if (pl->x16 && xp.hnblocks > 0) {
                // X list: blocks over whole classes, the reflection applied
                // to the multipoles here and to the results in k_v_reduce
                k_x_reflect_half<<<xp.npairs, 64, 0, st>>>(xp.A, xp.mask, pl->gtab, R4, KP, pl->Mh, pl->ebox, pl->Bx,
                                                           pl->ebx);
                gemm(xp, xp.hcls_off_d, pl->xiota, pl->Bx, pl->ebx, pl->xkeyslot,
                     pl->Yv + fp24_offset(vp.npairs, YS, 0), pl->y_rows - vp.npairs, false);
            }
            if (!pl->defer_reduce)
                k_v_reduce<R4><<<pl->n_m2l_work, R4 / 4 <= 128 ? 128 : 256, 0, st>>>(
                    pl->m2l_work, pl->ytptr, pl->ytlist, pl->Yv, vp.npairs, xp.mask, pl->stab, pl->L);
            IBIM_LAUNCH_CHECK();
            return IBIM_OK;
        }
    }
    if constexpr (P >= 7) {
        if (mixed && pl->prm.precision == IBIM_PREC_MIXED_F16 && pl->f16_ready && &vp == &pl->vp) {
            constexpr int KP = (R4 + 63) / 64 * 64;
            k_m_to_half<<<pl->nnodes, 64, 0, st>>>(pl->M, R4, KP, pl->Mh, pl->ebox);
            // V pairs, then (x16) the X pairs; tiles of 512 threads, at most
            // 2^22 per launch (32-bit work-item count)
            auto launch = [&](auto kern, const VPairs &pp, const int *key_off, const int *pA, const int *keyslot,
                              const _Float16 *B, const int *eb) {
                const int64_t total = (int64_t)pp.hnblocks * (pl->MRP / kHtRows);
                if (total >= (int64_t(1) << 31)) {
                    set_error("FP16 M2L: too many pair tiles for one plan");
                    return IBIM_ERR_STATE;
                }
                constexpr int64_t chunk = int64_t(1) << 22;
                for (int64_t first = 0; first < total; first += chunk) {
                    const int count = (int)std::min(chunk, total - first);
                    const int per_xcd = (count + 7) / 8;
                    kern<<<8 * per_xcd, 512, 0, st>>>(pp.hblk_cls, pp.hblk_j0, key_off, pA, pp.B, keyslot, pl->ops16,
                                                      pl->esig, pl->MRP, B, eb, pp.mask, pl->stab, pl->L, (int)first,
                                                      count, per_xcd);
                }
                return IBIM_OK;
            };
            int rc = launch(k_m2l_gemm_ht<P, false>, vp, vp.cls_off_d, vp.A, pl->keyslot, pl->Mh, pl->ebox);
            const VPairs &xp = pl->xp;
            if (!rc && pl->x16 && xp.hnblocks > 0) {
                k_x_reflect_half<<<xp.npairs, 64, 0, st>>>(xp.A, xp.mask, pl->gtab, R4, KP, pl->Mh, pl->ebox, pl->Bx,
                                                           pl->ebx);
                rc = launch(k_m2l_gemm_ht<P, true>, xp, xp.hcls_off_d, pl->xiota, pl->xkeyslot, pl->Bx, pl->ebx);
            }
            if (rc) return rc;
            IBIM_LAUNCH_CHECK();
            return IBIM_OK;
        }
    }
    if (mixed && pl->ops32) return launch_gemm<P, float>(vp, pl, pl->ops32, st);
    return launch_gemm<P, double>(vp, pl, pl->ops, st);
}

}  // namespace

// operators of every (level >= 2, |o|) class, FP64
int build_m2l_operators(ibim_plan *pl) {
    const int R4 = 4 * pl->r;
    pl->MRP = (R4 + 63) / 64 * 64;
    pl->KRP = pl->MRP;
    pl->n_cls = pl->nlevels > 2 ? (pl->nlevels - 2) * kNCanon : 0;
    if (pl->n_cls == 0) return IBIM_OK;
    int rc = dalloc(pl, &pl->ops, (size_t)pl->n_cls * pl->MRP * pl->KRP);
    if (rc) return rc;
    switch (pl->p) {
    case 2: return build_ops_p<2>(pl);
    case 3: return build_ops_p<3>(pl);
    case 4: return build_ops_p<4>(pl);
    case 5: return build_ops_p<5>(pl);
    case 6: return build_ops_p<6>(pl);
    case 7: return build_ops_p<7>(pl);
    case 8: return build_ops_p<8>(pl);
    }
    return IBIM_ERR_ARG;
}

// V pairs of the boxes flagged in `need`, grouped by operator class
namespace {
// first[k] = the first position of key k in the sorted keys (untouched
// where k is absent)
__global__ void k_key_starts(int np, const int *__restrict__ key, int *__restrict__ first) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < np && (j == 0 || key[j] != key[j - 1])) first[key[j]] = j;
}
__global__ void k_count_targets(int np, const int *__restrict__ B, int *__restrict__ cnt) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < np) atomicAdd(&cnt[B[j]], 1);
}
__global__ void k_iota_pairs(int np, int *__restrict__ v) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < np) v[j] = j;
}
}  // namespace

int build_vpairs(ibim_plan *pl, const unsigned char *need, VPairs &vp, std::vector<void *> *tmp_owner) {
    hipStream_t st = pl->stream;
    const int nn = pl->nnodes;
    auto alloc = [&](auto **p, size_t cnt) -> int {
        if (tmp_owner) {
            hipError_t e = dev_malloc((void **)p, (cnt ? cnt : 1) * sizeof(**p));
            if (e != hipSuccess) {
                set_error("hipMalloc (V pairs)");
                return IBIM_ERR_NOMEM;
            }
            tmp_owner->push_back((void *)*p);
            return IBIM_OK;
        }
        return dalloc(pl, p, cnt);
    };
    vp = VPairs{};
    if (pl->n_cls == 0) return IBIM_OK;
    ScratchScope scope(pl);  // temporaries below
    int *cnt = nullptr, *off = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &cnt, ((size_t)nn + 1) * sizeof(int)));
    IBIM_HIP_CHECK(smalloc(pl, &off, ((size_t)nn + 1) * sizeof(int)));
    IBIM_HIP_CHECK(hipMemsetAsync(cnt, 0, ((size_t)nn + 1) * sizeof(int), st));
    k_vpairs<0><<<nblk(nn), 256, 0, st>>>(nn, need, pl->loff[1], pl->litem[1], pl->flags, pl->level, pl->ctr,
                                          pl->d_lvl_rad, cnt, nullptr, nullptr, nullptr, nullptr, nullptr);
    IBIM_LAUNCH_CHECK();
    size_t bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, off, nn + 1, st));
    void *tmp = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &tmp, bytes ? bytes : 1));
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, off, nn + 1, st));
    int np = 0;
    IBIM_HIP_CHECK(hipMemcpyAsync(&np, off + nn, sizeof(int), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    vp.npairs = np;
    int rc;
    int *a0 = nullptr, *b0 = nullptr, *key0 = nullptr, *key1 = nullptr, *iota = nullptr, *order = nullptr;
    unsigned char *m0 = nullptr;
    if (np > 0) {
        IBIM_HIP_CHECK(smalloc(pl, &a0, (size_t)np * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &b0, (size_t)np * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &m0, (size_t)np));
        IBIM_HIP_CHECK(smalloc(pl, &key0, (size_t)np * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &key1, (size_t)np * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &iota, (size_t)np * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &order, (size_t)np * sizeof(int)));
        k_vpairs<1><<<nblk(nn), 256, 0, st>>>(nn, need, pl->loff[1], pl->litem[1], pl->flags, pl->level, pl->ctr,
                                              pl->d_lvl_rad, nullptr, off, a0, b0, m0, key0);
        hipcub::CountingInputIterator<int> it(0);
        // iota via a copy of the counting iterator
        IBIM_HIP_CHECK(hipcub::DeviceSelect::If(nullptr, bytes, it, iota, cnt, np,
                                                [] __device__(int) { return true; }, st));
        IBIM_HIP_CHECK(smalloc(pl, &tmp, bytes ? bytes : 1));
        IBIM_HIP_CHECK(hipcub::DeviceSelect::If(tmp, bytes, it, iota, cnt, np, [] __device__(int) { return true; },
                                                st));
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
        int endbit = 1;
        while ((1 << endbit) <= 8 * pl->n_cls) ++endbit;
        bytes = 0;
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key0, key1, iota, order, np, 0, endbit, st));
        IBIM_HIP_CHECK(smalloc(pl, &tmp, bytes ? bytes : 1));
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, key0, key1, iota, order, np, 0, endbit, st));
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
        if ((rc = alloc(&vp.A, np)) || (rc = alloc(&vp.B, np)) || (rc = alloc(&vp.mask, np))) return rc;
        k_permute_pairs<<<nblk(np), 256, 0, st>>>(np, order, a0, b0, m0, vp.A, vp.B, vp.mask);
        IBIM_LAUNCH_CHECK();
    }
    // class offsets (the first pair of each key of the sorted keys, found on
    // the device; a key without pairs takes the next key's) and the GEMM
    // block table, on the host
    const int nkeys = 8 * pl->n_cls;  // (class, reflection mask)
    vp.cls_off.assign(nkeys + 1, -1);
    if (np > 0) {
        int *kfirst = nullptr;
        IBIM_HIP_CHECK(smalloc(pl, &kfirst, (size_t)nkeys * sizeof(int)));
        IBIM_HIP_CHECK(hipMemsetAsync(kfirst, 0xff, (size_t)nkeys * sizeof(int), st));
        k_key_starts<<<nblk(np), 256, 0, st>>>(np, key1, kfirst);
        IBIM_LAUNCH_CHECK();
        IBIM_HIP_CHECK(hipMemcpyAsync(vp.cls_off.data(), kfirst, (size_t)nkeys * sizeof(int), hipMemcpyDeviceToHost,
                                      st));
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    vp.cls_off[nkeys] = np;
    for (int c = nkeys - 1; c >= 0; --c)
        if (vp.cls_off[c] < 0) vp.cls_off[c] = vp.cls_off[c + 1];
    std::vector<int> bc, bj;
    for (int c = 0; c < nkeys; ++c)
        for (int j = vp.cls_off[c]; j < vp.cls_off[c + 1]; j += 64) {
            bc.push_back(c);
            bj.push_back(j);
        }
    vp.nblocks = (int)bc.size();
    if ((rc = alloc(&vp.cls_off_d, nkeys + 1)) || (rc = alloc(&vp.blk_cls, std::max(1, vp.nblocks))) ||
        (rc = alloc(&vp.blk_j0, std::max(1, vp.nblocks))))
        return rc;
    IBIM_HIP_CHECK(hipMemcpyAsync(vp.cls_off_d, vp.cls_off.data(), (nkeys + 1) * sizeof(int),
                                  hipMemcpyHostToDevice, st));
    if (vp.nblocks > 0) {
        IBIM_HIP_CHECK(hipMemcpyAsync(vp.blk_cls, bc.data(), bc.size() * sizeof(int), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(vp.blk_j0, bj.data(), bj.size() * sizeof(int), hipMemcpyHostToDevice, st));
    }
    // FP16 GEMM block table (blocks of 32 kH32WavesV pairs for p <= 5, 64
    // for p = 6, kHtPairs above)
    {
        const int hnp = pl->p <= 5 ? 32 * kH32WavesV : (pl->p == 6 ? 64 : kHtPairs);
        std::vector<int> hc, hj;
        for (int c = 0; c < nkeys; ++c)
            for (int j = vp.cls_off[c]; j < vp.cls_off[c + 1]; j += hnp) {
                hc.push_back(c);
                hj.push_back(j);
            }
        vp.hnblocks = (int)hc.size();
        if ((rc = alloc(&vp.hblk_cls, std::max(1, vp.hnblocks))) || (rc = alloc(&vp.hblk_j0, std::max(1, vp.hnblocks))))
            return rc;
        if (vp.hnblocks > 0) {
            IBIM_HIP_CHECK(hipMemcpyAsync(vp.hblk_cls, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice, st));
            IBIM_HIP_CHECK(hipMemcpyAsync(vp.hblk_j0, hj.data(), hj.size() * sizeof(int), hipMemcpyHostToDevice, st));
        }
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    return IBIM_OK;
}

// ------------------------------------------------------ X list as GEMMs ---
// An X interaction (target b, coarser non-empty source leaf a, X list of b)
// evaluates K between b's and a's Chebyshev grids (kernelEvalChebyshev,
// bbfmm.h:456-479), like a V interaction but across levels.  Its geometry is
// fixed by the level of b, dl = level b - level a and the integer vector o =
// (c_a - c_b) / r_b: per axis d = r_b (o - x_t) + r_a x_s.  As for the V
// list, reflections fold o onto |o| (the mask records the negative
// components, and the same gather/scatter tables apply), pairs are grouped by
// (class, mask), and run through k_m2l_gemm with one operator per class.
namespace {
constexpr int kXOff = 8192;  // |o| < 8192 per axis

__global__ void k_xpairs(int pass, int nnodes, const unsigned char *__restrict__ need, const int *__restrict__ xoff,
                         const int *__restrict__ xit, const int *__restrict__ flags, const int *__restrict__ level,
                         const double *__restrict__ ctr, const double *__restrict__ lrad, int *__restrict__ cnt,
                         const int *__restrict__ off, int *__restrict__ pA, int *__restrict__ pB,
                         unsigned char *__restrict__ pm, unsigned long long *__restrict__ pkey, int *__restrict__ bad) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nnodes) return;
    int c = 0;
    if (b != 0 && need[b]) {
        const int lb = level[b];
        const double *rd = lrad + 3 * lb;
        for (int q = xoff[b]; q < xoff[b + 1]; ++q) {
            const int a = xit[q];
            if (flags[a] & kEmpty) continue;
            if (pass == 1) {
                // class key: level (5 bits), dl (6), |o| (3 x 13), then the mask (3)
                unsigned long long key = ((unsigned long long)lb << 45) | ((unsigned long long)(lb - level[a]) << 39);
                int m = 0;
                for (int d = 0; d < 3; ++d) {
                    const int o = (int)rint((ctr[3 * a + d] - ctr[3 * b + d]) / rd[d]);
                    if (o <= -kXOff || o >= kXOff) *bad = 1;
                    if (o < 0) m |= 1 << d;
                    key |= (unsigned long long)(abs(o) & (kXOff - 1)) << (26 - 13 * d);
                }
                const int j = off[b] + c;
                pA[j] = a;
                pB[j] = b;
                pm[j] = (unsigned char)m;
                pkey[j] = (key << 3) | (unsigned long long)m;
            }
            ++c;
        }
    }
    if (pass == 0) cnt[b] = c;
}

// operator entries of X class c: row g*r + t, column f*r + s
template <int P>
__global__ void k_build_xops(int ncls, const int *__restrict__ cls, const double *__restrict__ lrad, ChebTab tab,
                             PBConst kc, int MRP, int KRP, double *__restrict__ ops) {
    constexpr int R = P * P * P;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)R * R;
    if (e >= per * ncls) return;
    const int c = (int)(e / per);
    const int t = (int)((e % per) / R), s = (int)(e % R);
    const int lb = cls[5 * c], la = lb - cls[5 * c + 1];
    const double *rb = lrad + 3 * lb, *ra = lrad + 3 * la;
    const int ti[3] = {t % P, (t / P) % P, t / (P * P)};
    const int si[3] = {s % P, (s / P) % P, s / (P * P)};
    double d[3];
    for (int q = 0; q < 3; ++q) {
        double xs = 0.0, xt = 0.0;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            if (i == si[q]) xs = tab.xbar[i];
            if (i == ti[q]) xt = tab.xbar[i];
        }
        d[q] = rb[q] * (cls[5 * c + 2 + q] - xt) + ra[q] * xs;
    }
    double blk[4][4];
    for (int f = 0; f < 4; ++f) {
        double q[4] = {0, 0, 0, 0};
        q[f] = 1.0;
        double f0 = 0, fx = 0, fy = 0, fz = 0;
        pair_accum<double>(d[0], d[1], d[2], q[0], q[1], q[2], q[3], kc, f0, fx, fy, fz);
        blk[0][f] = f0;
        blk[1][f] = fx;
        blk[2][f] = fy;
        blk[3][f] = fz;
    }
    double *op = ops + (size_t)c * MRP * KRP;
    for (int g = 0; g < 4; ++g)
        for (int f = 0; f < 4; ++f) op[(size_t)(g * R + t) * KRP + f * R + s] = blk[g][f];
}

template <int P>
int build_xops_p(ibim_plan *pl, const int *d_cls) {
    constexpr int R = P * P * P;
    const int64_t total = (int64_t)R * R * pl->n_xcls;
    IBIM_HIP_CHECK(hipMemsetAsync(pl->xops, 0, (size_t)pl->n_xcls * pl->MRP * pl->KRP * sizeof(double), pl->stream));
    k_build_xops<P><<<nblk(total), 256, 0, pl->stream>>>(pl->n_xcls, d_cls, pl->d_lvl_rad, pl->tab, pl->kc, pl->MRP,
                                                         pl->KRP, pl->xops);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

template <int P>
int gemm_x_p(ibim_plan *pl, bool mixed) {
    const VPairs &xp = pl->xp;
    if (xp.nblocks == 0) return IBIM_OK;
    if (mixed && pl->prm.precision == IBIM_PREC_MIXED_F16 && pl->x16) return IBIM_OK;  // in the FP16 V stage
    if (mixed && pl->xops32) return launch_gemm<P, float>(xp, pl, pl->xops32, pl->stream);
    return launch_gemm<P, double>(xp, pl, pl->xops, pl->stream);
}
}  // namespace

// X pairs of the plan's own points, their classes and FP64 operators (run
// after the V operators, which also build the gather tables)
int build_xpairs(ibim_plan *pl) {
    pl->xp = VPairs{};
    pl->n_xcls = 0;
    if (!pl->ops || !pl->gtab) return IBIM_OK;
    hipStream_t st = pl->stream;
    const int nn = pl->nnodes;
    ScratchScope scope(pl);  // temporaries below
    auto salloc = [&](auto **p, size_t n) -> int {
        hipError_t e = smalloc(pl, p, (n ? n : 1) * sizeof(**p));
        if (e != hipSuccess) {
            set_error("hipMalloc (X pairs)");
            return IBIM_ERR_NOMEM;
        }
        return IBIM_OK;
    };
    int rc;
    int *cnt = nullptr, *off = nullptr, *bad = nullptr;
    if ((rc = salloc(&cnt, (size_t)nn + 1)) || (rc = salloc(&off, (size_t)nn + 1)) || (rc = salloc(&bad, 1))) return rc;
    IBIM_HIP_CHECK(hipMemsetAsync(cnt, 0, ((size_t)nn + 1) * sizeof(int), st));
    IBIM_HIP_CHECK(hipMemsetAsync(bad, 0, sizeof(int), st));
    k_xpairs<<<nblk(nn), 256, 0, st>>>(0, nn, pl->need_local, pl->loff[3], pl->litem[3], pl->flags, pl->level, pl->ctr,
                                       pl->d_lvl_rad, cnt, nullptr, nullptr, nullptr, nullptr, nullptr, bad);
    IBIM_LAUNCH_CHECK();
    size_t bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, off, nn + 1, st));
    void *tmp = nullptr;
    if ((rc = salloc((char **)&tmp, bytes))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, off, nn + 1, st));
    int np = 0;
    IBIM_HIP_CHECK(hipMemcpyAsync(&np, off + nn, sizeof(int), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    if (np == 0) return IBIM_OK;
    int *a0 = nullptr, *b0 = nullptr, *iota = nullptr, *order = nullptr;
    unsigned char *m0 = nullptr;
    unsigned long long *k0 = nullptr, *k1 = nullptr;
    if ((rc = salloc(&a0, np)) || (rc = salloc(&b0, np)) || (rc = salloc(&iota, np)) || (rc = salloc(&order, np)) ||
        (rc = salloc(&k0, np)) || (rc = salloc(&k1, np)) || (rc = salloc(&m0, np)))
        return rc;
    k_xpairs<<<nblk(nn), 256, 0, st>>>(1, nn, pl->need_local, pl->loff[3], pl->litem[3], pl->flags, pl->level, pl->ctr,
                                       pl->d_lvl_rad, nullptr, off, a0, b0, m0, k0, bad);
    k_iota_pairs<<<nblk(np), 256, 0, st>>>(np, iota);
    IBIM_LAUNCH_CHECK();
    bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k0, k1, iota, order, np, 0, 53, st));
    void *tmp2 = nullptr;
    if ((rc = salloc((char **)&tmp2, bytes))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp2, bytes, k0, k1, iota, order, np, 0, 53, st));
    int hbad = 0;
    std::vector<unsigned long long> hk(np);
    IBIM_HIP_CHECK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipMemcpyAsync(hk.data(), k1, (size_t)np * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    if (hbad) return IBIM_OK;  // offsets out of range: keep the direct X evaluation
    // classes: runs of equal class keys (the key without its 3 mask bits)
    std::vector<int> cls, cstart;
    for (int j = 0; j < np; ++j)
        if (j == 0 || (hk[j] >> 3) != (hk[j - 1] >> 3)) {
            const unsigned long long k = hk[j] >> 3;
            cls.push_back((int)(k >> 45));
            cls.push_back((int)((k >> 39) & 63));
            for (int d = 0; d < 3; ++d) cls.push_back((int)((k >> (26 - 13 * d)) & (kXOff - 1)));
            cstart.push_back(j);
        }
    const int nc = (int)cstart.size();
    VPairs &xp = pl->xp;
    xp.npairs = np;
    xp.cls_off.assign(8 * (size_t)nc + 1, 0);  // key 8 class + mask
    for (int c = 0, j = 0; c < nc; ++c)
        for (int k = 0; k < 8; ++k) {
            const int end = c + 1 < nc ? cstart[c + 1] : np;
            while (j < end && (int)(hk[j] & 7) == k) ++j;
            xp.cls_off[8 * c + k + 1] = j;
        }
    std::vector<int> bc, bj;
    for (int key = 0; key < 8 * nc; ++key)
        for (int j = xp.cls_off[key]; j < xp.cls_off[key + 1]; j += 64) {
            bc.push_back(key);
            bj.push_back(j);
        }
    xp.nblocks = (int)bc.size();
    // FP16 GEMM blocks over whole classes: that form applies the reflection
    // to the multipoles and to the results, so all 8 masks of a class share
    // the class operator
    std::vector<int> hc, hj, hco(nc + 1);
    for (int c = 0; c <= nc; ++c) hco[c] = xp.cls_off[8 * c];
    {
        const int hnp = pl->p <= 5 ? 32 * kH32WavesX : (pl->p == 6 ? 64 : kHtPairs);
        for (int c = 0; c < nc; ++c)
            for (int j = hco[c]; j < hco[c + 1]; j += hnp) {
                hc.push_back(c);
                hj.push_back(j);
            }
    }
    xp.hnblocks = (int)hc.size();
    int *d_cls = nullptr;
    if ((rc = dalloc(pl, &xp.A, np)) || (rc = dalloc(pl, &xp.B, np)) || (rc = dalloc(pl, &xp.mask, np)) ||
        (rc = dalloc(pl, &xp.cls_off_d, xp.cls_off.size())) || (rc = dalloc(pl, &xp.blk_cls, xp.nblocks)) ||
        (rc = dalloc(pl, &xp.blk_j0, xp.nblocks)) || (rc = dalloc(pl, &xp.hblk_cls, std::max(1, xp.hnblocks))) ||
        (rc = dalloc(pl, &xp.hblk_j0, std::max(1, xp.hnblocks))) || (rc = dalloc(pl, &xp.hcls_off_d, nc + 1)) ||
        (rc = salloc(&d_cls, cls.size())) ||
        (rc = dalloc(pl, &pl->xops, (size_t)nc * pl->MRP * pl->KRP)))
        return rc;
    k_permute_pairs<<<nblk(np), 256, 0, st>>>(np, order, a0, b0, m0, xp.A, xp.B, xp.mask);
    IBIM_LAUNCH_CHECK();
    IBIM_HIP_CHECK(hipMemcpyAsync(xp.cls_off_d, xp.cls_off.data(), xp.cls_off.size() * sizeof(int),
                                  hipMemcpyHostToDevice, st));
    IBIM_HIP_CHECK(hipMemcpyAsync(xp.hcls_off_d, hco.data(), hco.size() * sizeof(int), hipMemcpyHostToDevice, st));
    IBIM_HIP_CHECK(hipMemcpyAsync(xp.blk_cls, bc.data(), bc.size() * sizeof(int), hipMemcpyHostToDevice, st));
    IBIM_HIP_CHECK(hipMemcpyAsync(xp.blk_j0, bj.data(), bj.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (xp.hnblocks > 0) {
        IBIM_HIP_CHECK(hipMemcpyAsync(xp.hblk_cls, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(xp.hblk_j0, hj.data(), hj.size() * sizeof(int), hipMemcpyHostToDevice, st));
    }
    IBIM_HIP_CHECK(hipMemcpyAsync(d_cls, cls.data(), cls.size() * sizeof(int), hipMemcpyHostToDevice, st));
    pl->n_xcls = nc;
    switch (pl->p) {
    case 2: rc = build_xops_p<2>(pl, d_cls); break;
    case 3: rc = build_xops_p<3>(pl, d_cls); break;
    case 4: rc = build_xops_p<4>(pl, d_cls); break;
    case 5: rc = build_xops_p<5>(pl, d_cls); break;
    case 6: rc = build_xops_p<6>(pl, d_cls); break;
    case 7: rc = build_xops_p<7>(pl, d_cls); break;
    case 8: rc = build_xops_p<8>(pl, d_cls); break;
    default: rc = IBIM_ERR_ARG;
    }
    if (rc) return rc;
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    return IBIM_OK;
}

int m2l_gemm_x(ibim_plan *pl, bool mixed) {
    switch (pl->p) {
    case 2: return gemm_x_p<2>(pl, mixed);
    case 3: return gemm_x_p<3>(pl, mixed);
    case 4: return gemm_x_p<4>(pl, mixed);
    case 5: return gemm_x_p<5>(pl, mixed);
    case 6: return gemm_x_p<6>(pl, mixed);
    case 7: return gemm_x_p<7>(pl, mixed);
    case 8: return gemm_x_p<8>(pl, mixed);
    }
    return IBIM_ERR_ARG;
}

template <int P>
int reduce_p(ibim_plan *pl) {
    constexpr int R4 = 4 * P * P * P;
    if constexpr (P <= 6) {
        k_v_reduce<R4><<<pl->n_m2l_work, R4 / 4 <= 128 ? 128 : 256, 0, pl->stream>>>(
            pl->m2l_work, pl->ytptr, pl->ytlist, pl->Yv, pl->vp.npairs, pl->xp.mask, pl->stab, pl->L);
        IBIM_LAUNCH_CHECK();
    }
    return IBIM_OK;
}

int m2l_reduce(ibim_plan *pl) {
    switch (pl->p) {
    case 2: return reduce_p<2>(pl);
    case 3: return reduce_p<3>(pl);
    case 4: return reduce_p<4>(pl);
    case 5: return reduce_p<5>(pl);
    case 6: return reduce_p<6>(pl);
    }
    return IBIM_OK;
}

int m2l_gemm(ibim_plan *pl, const VPairs &vp, bool mixed) {
    switch (pl->p) {
    case 2: return gemm_p<2>(pl, vp, mixed);
    case 3: return gemm_p<3>(pl, vp, mixed);
    case 4: return gemm_p<4>(pl, vp, mixed);
    case 5: return gemm_p<5>(pl, vp, mixed);
    case 6: return gemm_p<6>(pl, vp, mixed);
    case 7: return gemm_p<7>(pl, vp, mixed);
    case 8: return gemm_p<8>(pl, vp, mixed);
    }
    return IBIM_ERR_ARG;
}

// FP16 operators, their row exponents and the per-matvec FP16 multipole and
// interaction-result buffers of the IBIM_PREC_MIXED_F16 matvec (the plan's V
// and X pairs must exist).  The X list joins the V list in this stage: its
// classes get (unreflected) operators after the V keys', its results the
// rows of Yv after the V pairs', and one per-target list over both (V pairs
// first, each kind in pair order) drives the reduction.
static int build_xops_f32(ibim_plan *pl);

int build_m2l_operators_f16(ibim_plan *pl) {
    if (pl->f16_ready || !pl->ops || pl->n_cls == 0) return IBIM_OK;
    hipStream_t st = pl->stream;
    const int R4 = 4 * pl->r;
    const int KP = (R4 + 63) / 64 * 64;
    const int YS = (R4 + 31) / 32 * 32;
    // p >= 7: both lists on k_m2l_gemm_ht, whose results go into L directly
    // (no pair rows, no per-target lists)
    const bool large = pl->p > 6;
    if (large && (pl->MRP % kHtRows != 0 || KP != pl->KRP)) {
        set_error("FP16 M2L: operator shape does not fit the p >= 7 tiles");
        return IBIM_ERR_STATE;
    }
    const VPairs &vp = pl->vp, &xp = pl->xp;
    const bool with_x = xp.npairs > 0 && pl->xops && xp.hnblocks > 0;
    const int npv = vp.npairs, npx = with_x ? xp.npairs : 0, npa = npv + npx;
    ScratchScope scope(pl);  // temporaries below
    auto salloc = [&](auto **p, size_t n) -> int {
        hipError_t e = smalloc(pl, p, (n ? n : 1) * sizeof(**p));
        if (e != hipSuccess) {
            set_error("hipMalloc (FP16 M2L set-up)");
            return IBIM_ERR_NOMEM;
        }
        return IBIM_OK;
    };
    int rc;
    if ((int64_t)vp.npairs + xp.npairs + 1 > (int64_t)0x7fffffff) {
        set_error("FP16 M2L: more than 2^31 - 2 V and X interactions");
        return IBIM_ERR_ARG;
    }
    if ((rc = dalloc(pl, &pl->Mh, (size_t)pl->nnodes * KP)) || (rc = dalloc(pl, &pl->ebox, (size_t)pl->nnodes)))
        return rc;
    if (!large) {
        if ((rc = dalloc(pl, &pl->Yv, fp24_buffer_dwords((int64_t)npa + 1, YS)))) return rc;
        pl->yv_dwords = fp24_buffer_dwords((int64_t)npa + 1, YS);
        pl->y_rows = npa;
    }
    // unfolded operators for the (class, mask) keys that have pairs: V keys
    // in slots [0, nsv), X keys in [nsv, nsv + nsx)
    auto slots = [](const std::vector<int> &off, int first, std::vector<int> &keyslot, std::vector<int> &slot_key) {
        keyslot.assign(off.size() > 0 ? off.size() - 1 : 0, -1);
        for (size_t k = 0; k + 1 < off.size(); ++k)
            if (off[k + 1] > off[k]) {
                keyslot[k] = first + (int)slot_key.size();
                slot_key.push_back((int)k);
            }
    };
    std::vector<int> keyslot, slot_key, xkeyslot, xunf;
    slots(vp.cls_off, 0, keyslot, slot_key);
    const int nsv = (int)slot_key.size();
    // X list: one slot per class (its unreflected operator, mask 0), indexed
    // by class; xunf is the same in the key-indexed form the unfolding takes
    if (with_x) {
        const int nc = (int)(xp.cls_off.size() - 1) / 8;
        xkeyslot.assign(nc, -1);
        xunf.assign(8 * (size_t)nc, -1);
        for (int c = 0; c < nc; ++c)
            if (xp.cls_off[8 * c + 8] > xp.cls_off[8 * c]) xunf[8 * c] = xkeyslot[c] = nsv + c;
    }
    const int nsx = (int)xkeyslot.size();
    const int ns = nsv + nsx;
    if ((rc = dalloc(pl, &pl->keyslot, std::max<size_t>(1, keyslot.size()))) ||
        (rc = dalloc(pl, &pl->xkeyslot, std::max<size_t>(1, xkeyslot.size()))) ||
        (rc = dalloc(pl, &pl->ops16, std::max<size_t>(1, (size_t)ns * pl->MRP * pl->KRP))) ||
        (rc = dalloc(pl, &pl->esig, std::max<size_t>(1, (size_t)ns * pl->MRP))))
        return rc;
    if (!keyslot.empty())
        IBIM_HIP_CHECK(hipMemcpyAsync(pl->keyslot, keyslot.data(), keyslot.size() * sizeof(int),
                                      hipMemcpyHostToDevice, st));
    if (!xkeyslot.empty())
        IBIM_HIP_CHECK(hipMemcpyAsync(pl->xkeyslot, xkeyslot.data(), xkeyslot.size() * sizeof(int),
                                      hipMemcpyHostToDevice, st));
    // per class with slots: its 8 (mask) slots, -1 where the key has no pairs
    auto unfold = [&](const std::vector<int> &ks, const double *ops) -> int {
        std::vector<int> cls, cslots;
        for (size_t k = 0; k < ks.size(); ++k) {
            if (ks[k] < 0) continue;
            const int c = (int)(k >> 3);
            if (cls.empty() || cls.back() != c) {
                cls.push_back(c);
                cslots.insert(cslots.end(), 8, -1);
            }
            cslots[8 * (cls.size() - 1) + (k & 7)] = ks[k];
        }
        if (cls.empty()) return IBIM_OK;
        int *d_cls = nullptr, *d_cslots = nullptr;
        IBIM_HIP_CHECK(smalloc(pl, &d_cls, cls.size() * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &d_cslots, cslots.size() * sizeof(int)));
        IBIM_HIP_CHECK(hipMemcpyAsync(d_cls, cls.data(), cls.size() * sizeof(int), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(
            hipMemcpyAsync(d_cslots, cslots.data(), cslots.size() * sizeof(int), hipMemcpyHostToDevice, st));
        k_unfold_half<<<dim3(pl->MRP, (unsigned)cls.size()), 64, pl->KRP * sizeof(double), st>>>(
            d_cls, d_cslots, ops, pl->MRP, pl->KRP, R4, pl->gtab, pl->stab, pl->ops16, pl->esig);
        IBIM_LAUNCH_CHECK();
        return IBIM_OK;
    };
    if ((rc = unfold(keyslot, pl->ops)) || (with_x && (rc = unfold(xunf, pl->xops)))) return rc;
    if (with_x) {
        if ((rc = dalloc(pl, &pl->Bx, (size_t)npx * KP)) || (rc = dalloc(pl, &pl->ebx, (size_t)npx)) ||
            (rc = dalloc(pl, &pl->xiota, (size_t)npx)))
            return rc;
        k_iota_pairs<<<nblk(npx), 256, 0, st>>>(npx, pl->xiota);
        IBIM_LAUNCH_CHECK();
    }
    if (large) {
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
        pl->x16 = nsx > 0;
        if (!pl->x16 && (rc = build_xops_f32(pl))) return rc;
        pl->f16_ready = true;
        return IBIM_OK;
    }
    // rows of Yv per target node: a stable sort of the target boxes of the
    // V pairs followed by the X pairs
    const int nn = pl->nnodes;
    int *bkey = nullptr, *bsorted = nullptr, *iota = nullptr, *cnt = nullptr;
    if ((rc = dalloc(pl, &pl->ytptr, (size_t)nn + 1)) || (rc = dalloc(pl, &pl->ytlist, std::max(1, npa))) ||
        (rc = salloc(&bkey, (size_t)npa)) || (rc = salloc(&bsorted, (size_t)npa)) || (rc = salloc(&iota, (size_t)npa)) ||
        (rc = salloc(&cnt, (size_t)nn + 1)))
        return rc;
    IBIM_HIP_CHECK(hipMemsetAsync(cnt, 0, ((size_t)nn + 1) * sizeof(int), st));
    if (npa > 0) {
        if (npv > 0) IBIM_HIP_CHECK(hipMemcpyAsync(bkey, vp.B, (size_t)npv * sizeof(int), hipMemcpyDeviceToDevice, st));
        if (npx > 0)
            IBIM_HIP_CHECK(hipMemcpyAsync(bkey + npv, xp.B, (size_t)npx * sizeof(int), hipMemcpyDeviceToDevice, st));
        k_count_targets<<<nblk(npa), 256, 0, st>>>(npa, bkey, cnt);
        k_iota_pairs<<<nblk(npa), 256, 0, st>>>(npa, iota);
        IBIM_LAUNCH_CHECK();
    }
    size_t bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, pl->ytptr, nn + 1, st));
    void *tmp = nullptr;
    if ((rc = salloc((char **)&tmp, bytes))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, pl->ytptr, nn + 1, st));
    if (npa > 0) {
        int endbit = 1;
        while ((1 << endbit) <= nn) ++endbit;
        bytes = 0;
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, bkey, bsorted, iota, pl->ytlist, npa, 0,
                                                          endbit, st));
        void *tmp2 = nullptr;
        if ((rc = salloc((char **)&tmp2, bytes))) return rc;
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp2, bytes, bkey, bsorted, iota, pl->ytlist, npa, 0, endbit,
                                                          st));
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    pl->x16 = nsx > 0;
    if (!pl->x16 && (rc = build_xops_f32(pl))) return rc;  // X list in FP32 GEMMs or direct
    // only a complete build enables the FP16 stage (a failed one leaves the
    // plan on the FP32 / FP64 GEMMs, and a retry rebuilds from the start)
    pl->f16_ready = true;
    return IBIM_OK;
}

// FP32 copy of the X-list operators (both mixed forms)
static int build_xops_f32(ibim_plan *pl) {
    if (pl->xops32 || !pl->xops || pl->n_xcls == 0) return IBIM_OK;
    const int64_t n = (int64_t)pl->n_xcls * pl->MRP * pl->KRP;
    int rc = dalloc(pl, &pl->xops32, (size_t)n);
    if (rc) return rc;
    k_to_float<<<nblk(n), 256, 0, pl->stream>>>(pl->xops, n, pl->xops32);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

// FP32 copy of the operators for the mixed-precision matvec
int build_m2l_operators_f32(ibim_plan *pl) {
    if (int rc = build_xops_f32(pl)) return rc;
    if (pl->ops32 || !pl->ops || pl->n_cls == 0) return IBIM_OK;
    const int64_t n = (int64_t)pl->n_cls * pl->MRP * pl->KRP;
    int rc = dalloc(pl, &pl->ops32, (size_t)n);
    if (rc) return rc;
    k_to_float<<<nblk(n), 256, 0, pl->stream>>>(pl->ops, n, pl->ops32);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

// ------------------------------------------------------- fused groups ---
// The M2L stage of a fused group matvec, in the order of the single-plan
// stage (FP64/FP32: X list, then V list, both into L by atomics; FP16:
// multipoles to FP16, V pairs, reflected X multipoles, X pairs, with the
// per-target reduction in group_m2l_reduce)
namespace {
template <int P>
int group_m2l_p(MatvecGroup *g) {
    constexpr int R4 = 4 * P * P * P;
    constexpr int KP = (R4 + 63) / 64 * 64;
    hipStream_t st = g->st;
    auto gl = [&](const MatvecGroup::Launch &L) {
        return GroupLaunch{g->d_mv, g->d_tab + L.off, L.aux >= 0 ? g->d_tab + L.aux : nullptr, g->k};
    };
    if (g->form == IBIM_PREC_MIXED_F16) {
        if constexpr (P <= 5) {
            if (g->m_to_half.blocks)
                k_m_to_half<true><<<g->m_to_half.blocks, 64, 0, st>>>(nullptr, R4, KP, nullptr, nullptr,
                                                                       gl(g->m_to_half));
            if (const int t = g->v_gemm.blocks) {
                const int pxcd = (t + 7) / 8;
                k_m2l_gemm_h32<P, kH32WavesV, true><<<8 * pxcd, 64 * kH32WavesV, 0, st>>>(
                    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr, 0, t,
                    pxcd, gl(g->v_gemm));
            }
            if (g->x_reflect.blocks)
                k_x_reflect_half<true><<<g->x_reflect.blocks, 64, 0, st>>>(nullptr, nullptr, nullptr, R4, KP, nullptr,
                                                                            nullptr, nullptr, nullptr, gl(g->x_reflect));
            if (const int t = g->x_gemm.blocks) {
                const int pxcd = (t + 7) / 8;
                k_m2l_gemm_h32<P, kH32WavesX, true><<<8 * pxcd, 64 * kH32WavesX, 0, st>>>(
                    nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr, 0, t,
                    pxcd, gl(g->x_gemm));
            }
            IBIM_LAUNCH_CHECK();
            return IBIM_OK;
        }
        set_error("matvec group: the FP16 form runs p <= 5");
        return IBIM_ERR_ARG;
    }
    auto run = [&](auto tag) {
        using T = decltype(tag);
        if (const int t = g->x_gemm.blocks) {
            const int pxcd = (t + 7) / 8;
            k_m2l_gemm<P, T, true, true><<<8 * pxcd, 256, 0, st>>>(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                   nullptr, nullptr, 0, 0, nullptr, nullptr, 0, t, pxcd,
                                                                   gl(g->x_gemm));
        }
        if (const int t = g->v_gemm.blocks) {
            const int pxcd = (t + 7) / 8;
            k_m2l_gemm<P, T, true, false><<<8 * pxcd, 256, 0, st>>>(nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                    nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, 0,
                                                                    t, pxcd, gl(g->v_gemm));
        }
    };
    if (g->form == IBIM_PREC_MIXED) run(float{});
    else run(double{});
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

template <int P>
int group_m2l_reduce_p(MatvecGroup *g) {
    constexpr int R4 = 4 * P * P * P;
    if constexpr (P <= 6) {
        if (g->v_reduce.blocks)
            k_v_reduce<R4, true><<<g->v_reduce.blocks, R4 / 4 <= 128 ? 128 : 256, 0, g->st>>>(
                nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr,
                GroupLaunch{g->d_mv, g->d_tab + g->v_reduce.off, nullptr, g->k});
        IBIM_LAUNCH_CHECK();
    }
    return IBIM_OK;
}
}  // namespace

int group_m2l(MatvecGroup *g) {
    switch (g->p) {
    case 2: return group_m2l_p<2>(g);
    case 3: return group_m2l_p<3>(g);
    case 4: return group_m2l_p<4>(g);
    case 5: return group_m2l_p<5>(g);
    case 6: return group_m2l_p<6>(g);
    }
    set_error("matvec group: unsupported Chebyshev order");
    return IBIM_ERR_ARG;
}

int group_m2l_reduce(MatvecGroup *g) {
    switch (g->p) {
    case 2: return group_m2l_reduce_p<2>(g);
    case 3: return group_m2l_reduce_p<3>(g);
    case 4: return group_m2l_reduce_p<4>(g);
    case 5: return group_m2l_reduce_p<5>(g);
    case 6: return group_m2l_reduce_p<6>(g);
    }
    set_error("matvec group: unsupported Chebyshev order");
    return IBIM_ERR_ARG;
}

}  // namespace ibim
