// ibim_fmm.hip -- the Chebyshev black-box FMM passes, all 16 PB kernels fused.
//
// Reference: bbfmm::kernel::upPass / downPass (bbfmm.h:500-623) run once per
// kernel closure (electric.cpp:352-367).  Every operator of that FMM is linear
// and shared by the 16 kernels, so one pass carries 4 multipole fields
// (the weighted charges qx, qy, qz, qpsi of electric.cpp:321-329) up the tree
// and 4 local fields (F0 = sum_a K1a[qa] - K2[qpsi], Fa = sum_b K3ab[qb] -
// K4a[qpsi]) down it.  That is the combination dmapping takes of the 16 FMM
// results (electric.cpp:369-384), reassociated.
//
//   P2M   leaf multipole M = S^T q, S the tensor Chebyshev interpolation
//         matrix at the standardized source coordinates (bbfmm.h:392-433,519)
//   M2M   M_parent += R[c]^T M_child, applied as three p x p contractions
//         instead of the dense r x r R[c] (bbfmm.h:328-367,530-534)
//   M2L   L_B += K(grid_B, grid_A) M_A over the V and X lists (bbfmm.h:547-561)
//   L2L   L_child += R[c] L_parent (bbfmm.h:566-567)
//   P2P   near field over the U and W lists, W boxes with all their subtree
//         sources (bbfmm.h:577-595)
//   L2P   pot += S L_leaf at the standardized target coordinate (bbfmm.h:600)
#include <cmath>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "ibim_internal.h"

namespace ibim {

namespace {

inline int nblk(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

// ----------------------------------------------------------------- P2M ---
// One block of BS threads per non-empty leaf, a thread per node (and field
// set); the leaf's points pass in chunks of BS, whose Chebyshev factors are
// formed one point per thread.  Each node sums its points in order.
// GROUP: one launch over the leaves of every member of a fused group (the
// group forms of the kernels below read their member's buffers from its
// view and take the same steps)
template <int P, int BS, bool GROUP = false>
__global__ __launch_bounds__(BS) void k_p2m(const int *__restrict__ leaves, const double *__restrict__ sx,
                                             const double *__restrict__ sy, const double *__restrict__ sz,
                                             const double *__restrict__ Q, int n, const int *__restrict__ sbeg,
                                             const int *__restrict__ send, const double *__restrict__ ctr,
                                             const int *__restrict__ level, const double *__restrict__ lrad,
                                             ChebTab tab, double *__restrict__ M, GroupLaunch g = {}) {
    constexpr int R = P * P * P;
    constexpr int PER = (R + BS - 1) / BS;
    constexpr int CH = BS;
    // rows padded by one double: lanes read P different rows at one point,
    // which an unpadded stride of CH doubles puts in one LDS bank
    __shared__ double sS[3][P][CH + 1];
    __shared__ double sq[4][CH];
    int bid = blockIdx.x;
    if constexpr (GROUP) {
        const MemberView &v = g.mv[group_member(g, blockIdx.x, bid)];
        leaves = uni(v.leaves_ne);
        sx = uni(v.sx);
        sy = uni(v.sy);
        sz = uni(v.sz);
        Q = uni(v.Q);
        n = uni(v.n);
        sbeg = uni(v.sbeg);
        send = uni(v.send);
        ctr = uni(v.ctr);
        level = uni(v.level);
        lrad = uni(v.lrad);
        M = uni(v.M);
    }
    const int b = gu<GROUP>(leaves[bid]);
    const int tid = threadIdx.x;
    const double cx = gu<GROUP>(ctr[3 * b]), cy = gu<GROUP>(ctr[3 * b + 1]), cz = gu<GROUP>(ctr[3 * b + 2]);
    const double *rd = lrad + 3 * gu<GROUP>(level[b]);
    const double rx = gu<GROUP>(rd[0]), ry = gu<GROUP>(rd[1]), rz = gu<GROUP>(rd[2]);
    double acc[PER][4];
#pragma unroll
    for (int k = 0; k < PER; ++k)
#pragma unroll
        for (int f = 0; f < 4; ++f) acc[k][f] = 0.0;
    const int beg = gu<GROUP>(sbeg[b]), end = gu<GROUP>(send[b]);
    for (int base = beg; base < end; base += CH) {
        const int cnt = min(CH, end - base);
        if (tid < cnt) {
            const int s = base + tid;
            double v[P];
            sp_values<P>((sx[s] - cx) / rx, tab, v);
#pragma unroll
            for (int m = 0; m < P; ++m) sS[0][m][tid] = v[m];
            sp_values<P>((sy[s] - cy) / ry, tab, v);
#pragma unroll
            for (int m = 0; m < P; ++m) sS[1][m][tid] = v[m];
            sp_values<P>((sz[s] - cz) / rz, tab, v);
#pragma unroll
            for (int m = 0; m < P; ++m) sS[2][m][tid] = v[m];
#pragma unroll
            for (int f = 0; f < 4; ++f) sq[f][tid] = Q[(size_t)f * n + s];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int m = tid + BS * k;
            if (m < R) {
                const int i = m % P, j = (m / P) % P, l = m / (P * P);
                for (int s = 0; s < cnt; ++s) {
                    const double w = sS[0][i][s] * sS[1][j][s] * sS[2][l][s];
                    acc[k][0] += w * sq[0][s];
                    acc[k][1] += w * sq[1][s];
                    acc[k][2] += w * sq[2][s];
                    acc[k][3] += w * sq[3][s];
                }
            }
        }
        __syncthreads();
    }
    double *Mb = M + (size_t)b * 4 * R;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int m = tid + BS * k;
        if (m < R)
#pragma unroll
            for (int f = 0; f < 4; ++f) Mb[f * R + m] = acc[k][f];
    }
}

// load the three p x p child transfer factors of child slot c into LDS
template <int P>
__device__ __forceinline__ void load_transfer(const ChebTab &tab, int c, double (*sT)[P][P]) {
    const int bx = c & 1, by = (c >> 1) & 1, bz = (c >> 2) & 1;
    for (int e = threadIdx.x; e < 3 * P * P; e += blockDim.x) {
        const int a = e / (P * P), i = (e / P) % P, l = e % P;
        const int b = a == 0 ? bx : (a == 1 ? by : bz);
        // kernel-argument arrays indexed at run time are read through a copy
        double v = 0.0;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int ii = 0; ii < P; ++ii)
#pragma unroll
                for (int ll = 0; ll < P; ++ll)
                    if (bb == b && ii == i && ll == l) v = tab.tr[bb][ii][ll];
        sT[a][i][l] = v;
    }
}

// ----------------------------------------------------------------- M2M ---
// out(f,l,m,n) += sum_i Tz(i,l) sum_j Ty(j,m) sum_k Tx(k,n) Mc(f,i,j,k)
// with T(i,l) = S_p(child node i, parent node l), index i*P*P + j*P + k.
// One block per parent; the children are transformed G at a time (all 8 on
// levels with few parents, where each block's chain of load latencies is
// the time), so their loads and the two inner contractions share one pass
// and three barriers; each output sums the children's terms in child order.
template <int P, int G, bool GROUP = false>
__global__ __launch_bounds__(256) void k_m2m(int lo, const int *__restrict__ flags, const int *__restrict__ fc,
                                             ChebTab tab, double *__restrict__ M, GroupLaunch g = {}) {
    constexpr int R = P * P * P;
    constexpr int R4 = 4 * R;
    constexpr int PER = (R4 + 255) / 256;
    static_assert(8 % G == 0, "children per pass");
    __shared__ double sIn[G * R4];
    __shared__ double sA[G * R4];
    __shared__ double sT[2][P][P];  // the transfer factors of the lower and the upper child half
    int id = lo + blockIdx.x;
    if constexpr (GROUP) {  // (aux: the member's first node of the level)
        int bid;
        const int m = group_member(g, blockIdx.x, bid);
        const MemberView &v = g.mv[m];
        id = uni(g.aux[m]) + bid;
        flags = uni(v.flags);
        fc = uni(v.first_child);
        M = uni(v.M);
    }
    if (gu<GROUP>(flags[id]) & kLeaf) return;
    const int tid = threadIdx.x;
    for (int e = tid; e < 2 * P * P; e += 256) {
        const int b = e / (P * P), i = (e / P) % P, l = e % P;
        // kernel-argument arrays indexed at run time are read through a copy
        double v = 0.0;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int ii = 0; ii < P; ++ii)
#pragma unroll
                for (int ll = 0; ll < P; ++ll)
                    if (bb == b && ii == i && ll == l) v = tab.tr[bb][ii][ll];
        sT[b][i][l] = v;
    }
    double acc[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) acc[k] = 0.0;
    const int c0 = gu<GROUP>(fc[id]);
    unsigned live = 0;  // bit c: child c is not empty (bbfmm.h:531 skips empty children)
    for (int c = 0; c < 8; ++c)
        if (!(gu<GROUP>(flags[c0 + c]) & kEmpty)) live |= 1u << c;
    for (int g0 = 0; g0 < 8; g0 += G) {
        if (!((live >> g0) & ((1u << G) - 1))) continue;  // (block-uniform)
        for (int e = tid; e < G * R4; e += 256) {
            const int c = e / R4;
            sIn[e] = (live >> (g0 + c)) & 1 ? M[(size_t)(c0 + g0 + c) * R4 + e % R4] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < G * R4; e += 256) {  // A(c,f,i,j,n)
            const int c = e / R4, rem = e % R, i = rem / (P * P), j = (rem / P) % P, nn = rem % P;
            const double(*T)[P] = sT[(g0 + c) & 1];
            const double *in = sIn + (e - rem) + i * P * P + j * P;
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < P; ++k) s += T[k][nn] * in[k];
            sA[e] = s;
        }
        __syncthreads();
        for (int e = tid; e < G * R4; e += 256) {  // B(c,f,i,m,n) -> sIn
            const int c = e / R4, rem = e % R, i = rem / (P * P), m = (rem / P) % P, nn = rem % P;
            const double(*T)[P] = sT[((g0 + c) >> 1) & 1];
            const double *in = sA + (e - rem) + i * P * P + nn;
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < P; ++j) s += T[j][m] * in[j * P];
            sIn[e] = s;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {  // out(f,l,m,n), children in order
            const int e = tid + 256 * k;
            if (e < R4) {
                const int f = e / R, rem = e % R, l = rem / (P * P), m = (rem / P) % P, nn = rem % P;
#pragma unroll
                for (int c = 0; c < G; ++c) {
                    if (!((live >> (g0 + c)) & 1)) continue;
                    const double(*T)[P] = sT[((g0 + c) >> 2) & 1];
                    const double *in = sIn + c * R4 + f * R + m * P + nn;
                    double s = 0.0;
#pragma unroll
                    for (int i = 0; i < P; ++i) s += T[i][l] * in[i * P * P];
                    acc[k] += s;
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = tid + 256 * k;
        if (e < R4) M[(size_t)id * R4 + e] = acc[k];
    }
}

// ----------------------------------------------------------------- L2L ---
// Lc(f,i,j,k) += sum_l Tz(i,l) sum_m Ty(j,m) sum_n Tx(k,n) Lp(f,l,m,n)
template <int P, bool GROUP = false>
__global__ __launch_bounds__(256) void k_l2l(int lo, const unsigned char *__restrict__ need,
                                             const int *__restrict__ parent, const int *__restrict__ slot,
                                             ChebTab tab, double *__restrict__ L, GroupLaunch g = {}) {
    constexpr int R = P * P * P;
    constexpr int R4 = 4 * R;
    __shared__ double sIn[R4];
    __shared__ double sA[R4];
    __shared__ double sT[3][P][P];
    int id = lo + blockIdx.x;
    if constexpr (GROUP) {  // (aux: the member's first node of the level)
        int bid;
        const int m = group_member(g, blockIdx.x, bid);
        const MemberView &v = g.mv[m];
        id = uni(g.aux[m]) + bid;
        need = uni(v.need_local);
        parent = uni(v.parent);
        slot = uni(v.slot);
        L = uni(v.L);
    }
    if (!gu<GROUP>((int)need[id])) return;
    const int tid = threadIdx.x;
    const int pid = gu<GROUP>(parent[id]);
    load_transfer<P>(tab, gu<GROUP>(slot[id]), sT);
    for (int e = tid; e < R4; e += 256) sIn[e] = L[(size_t)pid * R4 + e];
    __syncthreads();
    for (int e = tid; e < R4; e += 256) {  // A(f,l,m,k) = sum_n Tx(k,n) Lp(f,l,m,n)
        const int f = e / R, rem = e % R, l = rem / (P * P), m = (rem / P) % P, k = rem % P;
        const double *in = sIn + f * R + l * P * P + m * P;
        double s = 0.0;
#pragma unroll
        for (int nn = 0; nn < P; ++nn) s += sT[0][k][nn] * in[nn];
        sA[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < R4; e += 256) {  // B(f,l,j,k) = sum_m Ty(j,m) A(f,l,m,k)
        const int f = e / R, rem = e % R, l = rem / (P * P), j = (rem / P) % P, k = rem % P;
        const double *in = sA + f * R + l * P * P + k;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < P; ++m) s += sT[1][j][m] * in[m * P];
        sIn[e] = s;
    }
    __syncthreads();
    for (int e = tid; e < R4; e += 256) {  // C(f,i,j,k) = sum_l Tz(i,l) B(f,l,j,k)
        const int f = e / R, rem = e % R, i = rem / (P * P), j = (rem / P) % P, k = rem % P;
        const double *in = sIn + f * R + j * P + k;
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < P; ++l) s += sT[2][i][l] * in[l * P * P];
        L[(size_t)id * R4 + e] += s;
    }
}

// ------------------------------------------------------ M2L, direct form ---
// L_B = sum over non-empty A in V(B) and X(B) of K(grid_B, grid_A) M_A, the
// kernel matrix evaluated at the tensor Chebyshev grids of both boxes
// (kernelEvalChebyshev, bbfmm.h:456-479).  T = float (mixed precision):
// node coordinates relative to B's centre are formed in FP64 and rounded,
// the kernels are evaluated in FP32, and each source box's FP32 sums are
// accumulated in FP64.
template <int P, int BS, typename T, bool POLY>
__global__ __launch_bounds__(BS) void k_m2l_direct(int lists, const int *__restrict__ work, const int *__restrict__ voff,
                                                   const int *__restrict__ vit, const int *__restrict__ xoff,
                                                   const int *__restrict__ xit, const int *__restrict__ flags,
                                                   const double *__restrict__ ctr, const int *__restrict__ level,
                                                   const double *__restrict__ lrad, ChebTab tab, PBConst kc,
                                                   const double *__restrict__ M, double *__restrict__ L) {
    constexpr int R = P * P * P;
    constexpr int PER = (R + BS - 1) / BS;
    constexpr bool F64 = std::is_same<T, double>::value;
    __shared__ T spx[R], spy[R], spz[R];
    __shared__ T sm[4][R];
    __shared__ double sxb[P];
    const int id = work[blockIdx.x];
    const int tid = threadIdx.x;
    const bool useV = (lists & 1) && voff[id + 1] > voff[id];
    const bool useX = (lists & 2) && xoff[id + 1] > xoff[id];
    if (!useV && !useX) return;
    if (tid < P) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < P; ++i)
            if (i == tid) v = tab.xbar[i];
        sxb[tid] = v;
    }
    __syncthreads();
    const double *rb = lrad + 3 * level[id];
    // FP64: absolute coordinates, as the reference; FP32: relative to B's centre
    const double ox = F64 ? 0.0 : ctr[3 * id], oy = F64 ? 0.0 : ctr[3 * id + 1], oz = F64 ? 0.0 : ctr[3 * id + 2];
    T tx[PER], ty[PER], tz[PER];
    double acc[PER][4];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int m = tid + BS * k;
        const int i = m % P, j = (m / P) % P, l = (m / (P * P)) % P;
        tx[k] = (T)(ctr[3 * id] + rb[0] * sxb[i] - ox);
        ty[k] = (T)(ctr[3 * id + 1] + rb[1] * sxb[j] - oy);
        tz[k] = (T)(ctr[3 * id + 2] + rb[2] * sxb[l] - oz);
#pragma unroll
        for (int f = 0; f < 4; ++f) acc[k][f] = 0.0;
    }
    for (int lst = 0; lst < 2; ++lst) {
        if ((lst == 0 && !useV) || (lst == 1 && !useX)) continue;
        const int *off = lst == 0 ? voff : xoff;
        const int *it = lst == 0 ? vit : xit;
        for (int q = off[id]; q < off[id + 1]; ++q) {
            const int a = it[q];
            if (flags[a] & kEmpty) continue;
            const double *ra = lrad + 3 * level[a];
            const double ax = ctr[3 * a] - ox, ay = ctr[3 * a + 1] - oy, az = ctr[3 * a + 2] - oz;
            for (int e = tid; e < R; e += BS) {
                const int i = e % P, j = (e / P) % P, l = e / (P * P);
                spx[e] = (T)(ax + ra[0] * sxb[i]);
                spy[e] = (T)(ay + ra[1] * sxb[j]);
                spz[e] = (T)(az + ra[2] * sxb[l]);
            }
            const double *Ma = M + (size_t)a * 4 * R;
            for (int e = tid; e < 4 * R; e += BS) sm[e / R][e % R] = (T)Ma[e];
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (tid + BS * k < R) {
                    if constexpr (F64) {
                        double f0 = acc[k][0], fx = acc[k][1], fy = acc[k][2], fz = acc[k][3];
                        for (int s = 0; s < R; ++s)
                            pair_accum_m<POLY, double>(spx[s] - tx[k], spy[s] - ty[k], spz[s] - tz[k], sm[0][s],
                                                       sm[1][s], sm[2][s], sm[3][s], kc, f0, fx, fy, fz);
                        acc[k][0] = f0;
                        acc[k][1] = fx;
                        acc[k][2] = fy;
                        acc[k][3] = fz;
                    } else {
                        T f0 = 0, fx = 0, fy = 0, fz = 0;
                        for (int s = 0; s < R; ++s)
                            pair_accum_m<POLY, T>(spx[s] - tx[k], spy[s] - ty[k], spz[s] - tz[k], sm[0][s], sm[1][s],
                                                  sm[2][s], sm[3][s], kc, f0, fx, fy, fz);
                        acc[k][0] += (double)f0;
                        acc[k][1] += (double)fx;
                        acc[k][2] += (double)fy;
                        acc[k][3] += (double)fz;
                    }
                }
            }
            __syncthreads();
        }
    }
    double *Lb = L + (size_t)id * 4 * R;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int m = tid + BS * k;
        if (m < R)
#pragma unroll
            for (int f = 0; f < 4; ++f) Lb[f * R + m] += acc[k][f];
    }
}

// ----------------------------------------------------------- P2P + L2P ---
template <int P>
__global__ __launch_bounds__(64) void k_p2p_l2p(
    const int *__restrict__ tleaves, const int *__restrict__ tbeg, const int *__restrict__ tend,
    const double *__restrict__ tx, const double *__restrict__ ty, const double *__restrict__ tz,
    const double *__restrict__ sx, const double *__restrict__ sy, const double *__restrict__ sz,
    const double *__restrict__ Q, int ns, const int *__restrict__ uoff, const int *__restrict__ uit,
    const int *__restrict__ woff, const int *__restrict__ wit, const int *__restrict__ flags,
    const int *__restrict__ sbeg, const int *__restrict__ send, const double *__restrict__ ctr,
    const int *__restrict__ level, const double *__restrict__ lrad, ChebTab tab, PBConst kc,
    const double *__restrict__ L, double *__restrict__ F, int nt) {
    constexpr int R = P * P * P;
    constexpr int T = 64;
    __shared__ double s_x[T], s_y[T], s_z[T], s_q[4][T];
    __shared__ double sL[4 * R];
    const int b = tleaves[blockIdx.x];
    const int tid = threadIdx.x;
    for (int e = tid; e < 4 * R; e += T) sL[e] = L[(size_t)b * 4 * R + e];
    const double cx = ctr[3 * b], cy = ctr[3 * b + 1], cz = ctr[3 * b + 2];
    const double *rd = lrad + 3 * level[b];
    const int t_beg = tbeg[b], t_end = tend[b];
    for (int t0 = t_beg; t0 < t_end; t0 += T) {
        const int t = t0 + tid;
        const bool valid = t < t_end;
        const double px = valid ? tx[t] : 0.0, py = valid ? ty[t] : 0.0, pz = valid ? tz[t] : 0.0;
        double f0 = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        for (int lst = 0; lst < 2; ++lst) {
            const int *off = lst == 0 ? uoff : woff;
            const int *it = lst == 0 ? uit : wit;
            for (int q = off[b]; q < off[b + 1]; ++q) {
                const int a = it[q];
                if (flags[a] & kEmpty) continue;
                const int s_end = send[a];
                for (int s0 = sbeg[a]; s0 < s_end; s0 += T) {
                    const int cnt = min(T, s_end - s0);
                    __syncthreads();
                    if (tid < cnt) {
                        const int s = s0 + tid;
                        s_x[tid] = sx[s];
                        s_y[tid] = sy[s];
                        s_z[tid] = sz[s];
                        s_q[0][tid] = Q[s];
                        s_q[1][tid] = Q[(size_t)ns + s];
                        s_q[2][tid] = Q[2 * (size_t)ns + s];
                        s_q[3][tid] = Q[3 * (size_t)ns + s];
                    }
                    __syncthreads();
                    if (valid)
                        for (int s = 0; s < cnt; ++s)
                            pair_accum<double>(s_x[s] - px, s_y[s] - py, s_z[s] - pz, s_q[0][s], s_q[1][s],
                                               s_q[2][s], s_q[3][s], kc, f0, fx, fy, fz);
                }
            }
        }
        __syncthreads();  // sL loaded
        if (valid) {
            double Sx[P], Sy[P], Sz[P];
            sp_values<P>((px - cx) / rd[0], tab, Sx);
            sp_values<P>((py - cy) / rd[1], tab, Sy);
            sp_values<P>((pz - cz) / rd[2], tab, Sz);
            double g0 = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
            for (int l = 0; l < P; ++l)
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const double wyz = Sy[j] * Sz[l];
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        const int m = l * P * P + j * P + i;
                        const double w = Sx[i] * wyz;
                        g0 += w * sL[m];
                        gx += w * sL[R + m];
                        gy += w * sL[2 * R + m];
                        gz += w * sL[3 * R + m];
                    }
                }
            F[t] = f0 + g0;
            F[(size_t)nt + t] = fx + gx;
            F[2 * (size_t)nt + t] = fy + gy;
            F[3 * (size_t)nt + t] = fz + gz;
        }
    }
}


// ---------------------------------------------- near-field source ranges ---
// For each target leaf: the sources of its non-empty U and W boxes as ranges
// of the sorted point order, sorted and merged where they touch (Morton
// order makes neighbouring boxes' ranges contiguous).
template <int PASS>
__global__ __launch_bounds__(64) void k_p2p_ranges(int nleaves, const int *__restrict__ tleaves,
                                                   const int *__restrict__ uoff, const int *__restrict__ uit,
                                                   const int *__restrict__ woff, const int *__restrict__ wit,
                                                   const int *__restrict__ flags, const int *__restrict__ sbeg,
                                                   const int *__restrict__ send, int *__restrict__ cnt,
                                                   const int *__restrict__ roff, int2 *__restrict__ ranges) {
    // one wave per target leaf: collect the non-empty U and W boxes' source
    // ranges (ballot compaction), sort them by begin (rank sort: begins of
    // disjoint non-empty ranges are distinct) and merge touching ones
    constexpr int CAP = 512;
    __shared__ int2 loc[CAP], srt[CAP];
    __shared__ int runs[CAP + 1];
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int b = tleaves[i];
    int n = 0;
    for (int lst = 0; lst < 2; ++lst) {
        const int *off = lst == 0 ? uoff : woff;
        const int *it = lst == 0 ? uit : wit;
        for (int q0 = off[b]; q0 < off[b + 1]; q0 += 64) {
            const int q = q0 + lane;
            const int a = q < off[b + 1] ? it[q] : -1;
            const bool v = a >= 0 && !(flags[a] & kEmpty);
            const unsigned long long m = __ballot(v);
            const int pos = n + __popcll(m & ((1ull << lane) - 1ull));
            if (v && pos < CAP) loc[pos] = make_int2(sbeg[a], send[a]);
            n += __popcll(m);
        }
    }
    if (n > CAP) {  // keep the per-box kernel for this target set
        if (PASS == 0 && lane == 0) cnt[i] = -1;
        return;
    }
    __syncthreads();
    for (int e = lane; e < n; e += 64) {
        const int x = loc[e].x;
        int r = 0;
        for (int f = 0; f < n; ++f) r += loc[f].x < x;
        srt[r] = loc[e];
    }
    __syncthreads();
    // element k starts a run unless it continues the previous range
    int nr = 0;
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int k = k0 + lane;
        const bool st = k < n && (k == 0 || srt[k - 1].y != srt[k].x);
        const unsigned long long m = __ballot(st);
        if (st) runs[nr + __popcll(m & ((1ull << lane) - 1ull))] = k;
        nr += __popcll(m);
    }
    if (lane == 0) runs[nr] = n;
    __syncthreads();
    if (PASS == 0) {
        if (lane == 0) cnt[i] = nr;
    } else {
        for (int r = lane; r < nr; r += 64)
            ranges[roff[i] + r] = make_int2(srt[runs[r]].x, srt[runs[r + 1] - 1].y);
    }
}

// ------------------------------------------------------- P2P work units ---
// The near field of a target leaf is its merged source ranges (U list and
// the W list's subtree sources).  Those lists differ by orders of magnitude
// between leaves, so one block per leaf leaves a long tail of a few busy
// blocks.  The work is therefore cut into units of at most 64 targets x
// kP2PChunk sources, launched heaviest first; a leaf whose sources span
// several units writes one partial field per unit, and k_l2p sums them in
// unit order (deterministic).
constexpr int kP2PChunk = 1024;

template <typename T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<double> { using type = double4; };

// Each lane of the 64-thread block is (target mt, source slice sl):
// nsl = 64 / ntc slices of ntc targets, so 64 - nsl*ntc lanes idle.  Sources
// are staged in LDS 64 at a time (one per lane).  MIXED: coordinates relative
// to the unit's first target are formed in FP64 and rounded to FP32, the
// kernels are evaluated in FP32 (PB kernels by pair_pb32), each lane's FP32
// sums over one LDS batch (at most 64 terms) are accumulated in FP64.
// (GROUP: the units [aux[m], aux[m] + blocks) of each member m, whose own
// points are the targets and the sources)
template <bool MIXED, bool POLY, bool GROUP = false>
__global__ __launch_bounds__(64) void k_p2p_unit(
    const int4 *__restrict__ ua, const int4 *__restrict__ ub, const double *__restrict__ tx,
    const double *__restrict__ ty, const double *__restrict__ tz, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz, const double *__restrict__ Q, int ns,
    const int2 *__restrict__ ranges, PBConst kc, double *__restrict__ F, int nt, double *__restrict__ part,
    GroupLaunch g = {}) {
    using T = typename std::conditional<MIXED, float, double>::type;
    using T4 = typename Vec4<T>::type;
    __shared__ T4 sbuf[2][64];  // {x, y, z, q_psi}, {q_x, q_y, q_z, q_psi / (4 pi)} per source
    T4 *s0 = sbuf[0], *s1 = sbuf[1];
    __shared__ double red[4][64];
    const int lane = threadIdx.x;
    int u = blockIdx.x;
    if constexpr (GROUP) {
        int bid;
        const int m = group_member(g, blockIdx.x, bid);
        const MemberView &v = g.mv[m];
        u = uni(g.aux[m]) + bid;
        ua = uni(v.ua);
        ub = uni(v.ub);
        tx = sx = uni(v.sx);
        ty = sy = uni(v.sy);
        tz = sz = uni(v.sz);
        Q = uni(v.Q);
        ns = nt = uni(v.n);
        ranges = uni(v.ranges);
        kc = v.kc;
        F = uni(v.F);
        part = uni(v.part);
    }
    const int4 a = gu<GROUP>(ua[u]), c = gu<GROUP>(ub[u]);
    const int t0 = a.y, ntc = a.z, slot = a.w, total = c.z;
    const int nsl = 64 / ntc;
    const int mt = lane % ntc, sl = lane / ntc;
    const bool active = sl < nsl;
    const int t = t0 + mt;
    const double ox = MIXED ? gu<GROUP>(tx[t0]) : 0.0, oy = MIXED ? gu<GROUP>(ty[t0]) : 0.0,
                 oz = MIXED ? gu<GROUP>(tz[t0]) : 0.0;
    const T px = (T)(tx[t] - ox), py = (T)(ty[t] - oy), pz = (T)(tz[t] - oz);
    const PBConst32 k32 = pb_const32(kc);
    double f0 = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    int cr = c.x, cbase = -c.y;  // this lane's loading cursor: source j sits at ranges[cr].x + j - cbase
    for (int b0 = 0; b0 < total; b0 += 64) {
        const int j = b0 + lane;
        if (j < total) {
            while (j - cbase >= ranges[cr].y - ranges[cr].x) {
                cbase += ranges[cr].y - ranges[cr].x;
                ++cr;
            }
            const int s = ranges[cr].x + (j - cbase);
            T4 u, v;
            u.x = (T)(sx[s] - ox);
            u.y = (T)(sy[s] - oy);
            u.z = (T)(sz[s] - oz);
            u.w = (T)Q[3 * (size_t)ns + s];
            v.x = (T)Q[s];
            v.y = (T)Q[(size_t)ns + s];
            v.z = (T)Q[2 * (size_t)ns + s];
            v.w = (T)(Q[3 * (size_t)ns + s] * kc.inv4pi);  // qs of pair_pb32
            s0[lane] = u;
            s1[lane] = v;
        }
        __syncthreads();
        const int cnt = min(64, total - b0);
        if constexpr (MIXED && !POLY) {
          if (active) {
            // FP32 PB kernels: F0 = f0a / (4 pi) - f0b
            float ga = 0, gb = 0, gx = 0, gy = 0, gz = 0, ha = 0, hb = 0, hx = 0, hy = 0, hz = 0;
            // sources k = sl, sl + nsl, ... two per step; u at s0[k], v at
            // s1[k] = s0[k + 64]
            const float4 *pu = (const float4 *)s0 + sl;
            if (nsl == 1) {
                // (one source slice, the common case: a block-uniform
                // counter and constant LDS offsets, no per-lane loop control)
                int q = 0;
#pragma unroll 1
                for (; q + 1 < cnt; q += 2) {
                    const float4 u = s0[q], v = s1[q], u2 = s0[q + 1], v2 = s1[q + 1];
                    pair_pb32(u.x - (float)px, u.y - (float)py, u.z - (float)pz, v.x, v.y, v.z, v.w, k32, ga, gb,
                              gx, gy, gz);
                    pair_pb32(u2.x - (float)px, u2.y - (float)py, u2.z - (float)pz, v2.x, v2.y, v2.z, v2.w, k32, ha,
                              hb, hx, hy, hz);
                }
                pu += q;
            } else {
                const float4 *pend = (const float4 *)s0 + max(cnt - nsl, 0);
                for (; pu < pend; pu += 2 * nsl) {
                    const float4 u = pu[0], v = pu[64], u2 = pu[nsl], v2 = pu[nsl + 64];
                    pair_pb32(u.x - (float)px, u.y - (float)py, u.z - (float)pz, v.x, v.y, v.z, v.w, k32, ga, gb,
                              gx, gy, gz);
                    pair_pb32(u2.x - (float)px, u2.y - (float)py, u2.z - (float)pz, v2.x, v2.y, v2.z, v2.w, k32, ha,
                              hb, hx, hy, hz);
                }
            }
            const int k = (int)(pu - (const float4 *)s0);
            if (k < cnt) {
                const float4 u = s0[k], v = s1[k];
                pair_pb32(u.x - (float)px, u.y - (float)py, u.z - (float)pz, v.x, v.y, v.z, v.w, k32, ga, gb, gx, gy,
                          gz);
            }
            f0 += (double)(ga + ha) * kc.inv4pi - (double)(gb + hb);
            fx += (double)(gx + hx);
            fy += (double)(gy + hy);
            fz += (double)(gz + hz);
          }
        } else if constexpr (!MIXED && !POLY) {
          if (active) {
            // FP64 PB kernels: F0 = f0a / (4 pi) - f0b
            const double k2r = kc.k2reg / kc.inv4pi;
            double ga = 0, gb = 0, gx = 0, gy = 0, gz = 0, ha = 0, hb = 0, hx = 0, hy = 0, hz = 0;
            const double4 *pu = (const double4 *)s0 + sl;
            if (nsl == 1) {  // (as in the FP32 form)
                int q = 0;
#pragma unroll 2
                for (; q + 1 < cnt; q += 2) {
                    const double4 u = s0[q], v = s1[q], u2 = s0[q + 1], v2 = s1[q + 1];
                    pair_pb64(u.x - px, u.y - py, u.z - pz, v.x, v.y, v.z, v.w, kc, k2r, ga, gb, gx, gy, gz);
                    pair_pb64(u2.x - px, u2.y - py, u2.z - pz, v2.x, v2.y, v2.z, v2.w, kc, k2r, ha, hb, hx, hy, hz);
                }
                pu += q;
            } else {
                const double4 *pend = (const double4 *)s0 + max(cnt - nsl, 0);
                for (; pu < pend; pu += 2 * nsl) {
                    const double4 u = pu[0], v = pu[64], u2 = pu[nsl], v2 = pu[nsl + 64];
                    pair_pb64(u.x - px, u.y - py, u.z - pz, v.x, v.y, v.z, v.w, kc, k2r, ga, gb, gx, gy, gz);
                    pair_pb64(u2.x - px, u2.y - py, u2.z - pz, v2.x, v2.y, v2.z, v2.w, kc, k2r, ha, hb, hx, hy, hz);
                }
            }
            const int k = (int)(pu - (const double4 *)s0);
            if (k < cnt) {
                const double4 u = s0[k], v = s1[k];
                pair_pb64(u.x - px, u.y - py, u.z - pz, v.x, v.y, v.z, v.w, kc, k2r, ga, gb, gx, gy, gz);
            }
            f0 += (ga + ha) * kc.inv4pi - (gb + hb);
            fx += gx + hx;
            fy += gy + hy;
            fz += gz + hz;
          }
        } else if (active) {
            T g0 = 0, gx = 0, gy = 0, gz = 0, h0 = 0, hx = 0, hy = 0, hz = 0;
            int k = sl;
            for (; k + nsl < cnt; k += 2 * nsl) {
                const T4 u = s0[k], v = s1[k], u2 = s0[k + nsl], v2 = s1[k + nsl];
                pair_accum_m<POLY, T>(u.x - px, u.y - py, u.z - pz, v.x, v.y, v.z, u.w, kc, g0, gx, gy, gz);
                pair_accum_m<POLY, T>(u2.x - px, u2.y - py, u2.z - pz, v2.x, v2.y, v2.z, u2.w, kc, h0, hx, hy, hz);
            }
            if (k < cnt) {
                const T4 u = s0[k], v = s1[k];
                pair_accum_m<POLY, T>(u.x - px, u.y - py, u.z - pz, v.x, v.y, v.z, u.w, kc, g0, gx, gy, gz);
            }
            f0 += (double)(g0 + h0);
            fx += (double)(gx + hx);
            fy += (double)(gy + hy);
            fz += (double)(gz + hz);
        }
        __syncthreads();
    }
    red[0][lane] = f0;
    red[1][lane] = fx;
    red[2][lane] = fy;
    red[3][lane] = fz;
    __syncthreads();
    if (lane < ntc) {
        double v[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double acc = red[f][lane];
            for (int q = 1; q < nsl; ++q) acc += red[f][q * ntc + lane];
            v[f] = acc;
        }
        if (slot < 0) {
#pragma unroll
            for (int f = 0; f < 4; ++f) F[f * (size_t)nt + t] = v[f];
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f) part[(size_t)slot * 256 + f * 64 + c.w + lane] = v[f];
        }
    }
}

__device__ __forceinline__ int range_sources(const int *roff, const int2 *ranges, int i) {
    int ns = 0;
    for (int r = roff[i]; r < roff[i + 1]; ++r) ns += ranges[r].y - ranges[r].x;
    return ns;
}

// Target chunks of a leaf: full chunks of 64, then the remainder r as one
// chunk if r <= 32, else as 32 + (r - 32), so that a unit's lanes (targets x
// floor(64 / targets) source slices) stay mostly busy.  A leaf's partial
// fields are laid out by 64-target blocks; a remainder chunk's lanes start at
// offset 0 or 32 of its block.
__device__ __forceinline__ int p2p_target_chunks(int nt) {
    const int q = nt / 64, r = nt % 64;
    return q + (r == 0 ? 0 : (r > 32 ? 2 : 1));
}
__device__ __forceinline__ void p2p_target_chunk(int nt, int tc, int &first, int &count) {
    const int q = nt / 64, r = nt % 64;
    if (tc < q) {
        first = 64 * tc;
        count = 64;
    } else if (r > 32) {
        first = 64 * q + (tc == q ? 0 : 32);
        count = tc == q ? 32 : r - 32;
    } else {
        first = 64 * q;
        count = r;
    }
}

// units and partial slots per target leaf
__global__ void k_p2p_count(int nl, const int *__restrict__ tleaves, const int *__restrict__ tbeg,
                            const int *__restrict__ tend, const int *__restrict__ roff,
                            const int2 *__restrict__ ranges, int *__restrict__ ucnt, int *__restrict__ pcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const int b = tleaves[i];
    const int ns = range_sources(roff, ranges, i);
    const int nt = tend[b] - tbeg[b];
    const int ncs = max(1, (ns + kP2PChunk - 1) / kP2PChunk);
    ucnt[i] = p2p_target_chunks(nt) * ncs;
    pcnt[i] = ncs > 1 ? (nt + 63) / 64 * ncs : 0;
}

// unit records, their cost (per-lane pair evaluations), and the per-leaf
// split table read by k_l2p
__global__ void k_p2p_fill(int nl, const int *__restrict__ tleaves, const int *__restrict__ tbeg,
                           const int *__restrict__ tend, const int *__restrict__ roff,
                           const int2 *__restrict__ ranges, const int *__restrict__ uoff,
                           const int *__restrict__ poff, int4 *__restrict__ ua, int4 *__restrict__ ub,
                           int *__restrict__ cost, int *__restrict__ idx, int2 *__restrict__ lsplit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const int b = tleaves[i];
    const int ns = range_sources(roff, ranges, i);
    const int ntl = tend[b] - tbeg[b];
    const int nct = p2p_target_chunks(ntl);
    const int ncs = max(1, (ns + kP2PChunk - 1) / kP2PChunk);
    lsplit[i] = make_int2(ncs > 1 ? poff[i] : -1, ncs);
    int u = uoff[i];
    int r = roff[i], off = 0;  // cursor at the first source of the chunk
    for (int sc = 0; sc < ncs; ++sc) {
        const int cnt = max(0, min(kP2PChunk, ns - sc * kP2PChunk));
        for (int tc = 0; tc < nct; ++tc) {
            int first, ntc;
            p2p_target_chunk(ntl, tc, first, ntc);
            const int blk = first / 64;  // 64-target block of the partial-field layout
            ua[u] = make_int4(i, tbeg[b] + first, ntc, ncs > 1 ? poff[i] + blk * ncs + sc : -1);
            ub[u] = make_int4(r, off, cnt, first % 64);
            cost[u] = cnt / (64 / ntc) + 1;
            idx[u] = u;
            ++u;
        }
        int adv = cnt;
        while (adv > 0) {
            const int left = ranges[r].y - ranges[r].x - off;
            if (adv < left) {
                off += adv;
                adv = 0;
            } else {
                adv -= left;
                ++r;
                off = 0;
            }
        }
    }
}

// target x source pairs of each work unit (for the roofline accounting)
__global__ void k_unit_pairs(int n, const int4 *__restrict__ ua, const int4 *__restrict__ ub,
                             long long *__restrict__ out) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) out[u] = (long long)ua[u].z * ub[u].z;
}

__global__ void k_p2p_order(int n, const int *__restrict__ idx, const int4 *__restrict__ ua,
                            const int4 *__restrict__ ub, int4 *__restrict__ oa, int4 *__restrict__ ob) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    oa[k] = ua[idx[k]];
    ob[k] = ub[idx[k]];
}

// L2P (bbfmm.h:600): F += S L_leaf at each target, one block (one wave) per
// target leaf.  For a leaf whose near
// field was split over several work units (lsplit[leaf].y > 1) the near
// field is first summed from the units' partial fields, in unit order.
// (GROUP: the non-empty leaves of every member; the leaf's expansion is
// staged in LDS, since through a pointer read from memory its uniform reads
// would be vector loads)
template <int P, bool GROUP = false>
__global__ __launch_bounds__(64) void k_l2p(const int *__restrict__ tleaves, const int *__restrict__ tbeg,
                                            const int *__restrict__ tend, const double *__restrict__ tx,
                                            const double *__restrict__ ty, const double *__restrict__ tz,
                                            const double *__restrict__ ctr, const int *__restrict__ level,
                                            const double *__restrict__ lrad, ChebTab tab,
                                            const double *__restrict__ L, const int2 *__restrict__ lsplit,
                                            const double *__restrict__ part, double *__restrict__ F, int nt,
                                            GroupLaunch g = {}) {
    constexpr int R = P * P * P;
    __shared__ double sLg[GROUP ? 4 * R : 1];
    int bid = blockIdx.x;
    if constexpr (GROUP) {
        const MemberView &v = g.mv[group_member(g, blockIdx.x, bid)];
        tleaves = uni(v.leaves_ne);
        tbeg = uni(v.sbeg);
        tend = uni(v.send);
        tx = uni(v.sx);
        ty = uni(v.sy);
        tz = uni(v.sz);
        ctr = uni(v.ctr);
        level = uni(v.level);
        lrad = uni(v.lrad);
        L = uni(v.L);
        lsplit = uni(v.lsplit);
        part = uni(v.part);
        F = uni(v.F);
        nt = uni(v.n);
    }
    const int b = gu<GROUP>(tleaves[bid]);
    const int2 sp = lsplit ? gu<GROUP>(lsplit[bid]) : make_int2(-1, 1);
    // the leaf's expansion is read at wave-uniform addresses: scalar loads
    const double *sL = L + (size_t)b * 4 * R;
    if constexpr (GROUP) {
        for (int e = threadIdx.x; e < 4 * R; e += 64) sLg[e] = sL[e];
        __syncthreads();
        sL = sLg;
    }
    const double cx = gu<GROUP>(ctr[3 * b]), cy = gu<GROUP>(ctr[3 * b + 1]), cz = gu<GROUP>(ctr[3 * b + 2]);
    const double *rd = lrad + 3 * gu<GROUP>(level[b]);
    const int t_end = gu<GROUP>(tend[b]);
    for (int t = gu<GROUP>(tbeg[b]) + threadIdx.x; t < t_end; t += 64) {
        double Sx[P], Sy[P], Sz[P];
        sp_values<P>((tx[t] - cx) / rd[0], tab, Sx);
        sp_values<P>((ty[t] - cy) / rd[1], tab, Sy);
        sp_values<P>((tz[t] - cz) / rd[2], tab, Sz);
        double g0 = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
        if constexpr (GROUP) {
            // (one z-layer of the expansion per pass, rolled: unrolled, the
            // compiler reads all 4 r values ahead and spills)
#pragma unroll 1
            for (int l = 0; l < P; ++l) {
                double szl = Sz[0];
#pragma unroll
                for (int q = 1; q < P; ++q)
                    if (q == l) szl = Sz[q];
                const double *sl = sL + l * P * P;
#pragma unroll
                for (int jj = 0; jj < P; ++jj) {
                    const double wyz = Sy[jj] * szl;
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        const int m = jj * P + i;
                        const double w = Sx[i] * wyz;
                        g0 += w * sl[m];
                        gx += w * sl[R + m];
                        gy += w * sl[2 * R + m];
                        gz += w * sl[3 * R + m];
                    }
                }
            }
        } else {
        for (int l = 0; l < P; ++l)
#pragma unroll
            for (int jj = 0; jj < P; ++jj) {
                const double wyz = Sy[jj] * Sz[l];
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const int m = l * P * P + jj * P + i;
                    const double w = Sx[i] * wyz;
                    g0 += w * sL[m];
                    gx += w * sL[R + m];
                    gy += w * sL[2 * R + m];
                    gz += w * sL[3 * R + m];
                }
            }
        }
        double n0, nx, ny, nz;  // near field
        if (sp.y > 1) {
            const int i = t - tbeg[b];
            n0 = nx = ny = nz = 0.0;
            for (int sc = 0; sc < sp.y; ++sc) {
                const double *pp = part + (size_t)(sp.x + (i >> 6) * sp.y + sc) * 256 + (i & 63);
                n0 += pp[0];
                nx += pp[64];
                ny += pp[128];
                nz += pp[192];
            }
        } else {
            n0 = F[t];
            nx = F[(size_t)nt + t];
            ny = F[2 * (size_t)nt + t];
            nz = F[3 * (size_t)nt + t];
        }
        F[t] = n0 + g0;
        F[(size_t)nt + t] = nx + gx;
        F[2 * (size_t)nt + t] = ny + gy;
        F[3 * (size_t)nt + t] = nz + gz;
    }
}

// ------------------------------------------------------------- helpers ---
__global__ void k_charges(int n, const int *__restrict__ perm, const double *__restrict__ snx,
                          const double *__restrict__ sny, const double *__restrict__ snz,
                          const double *__restrict__ sw, const double *__restrict__ x, double *__restrict__ Q) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    const double phi = x[o], psi = x[(size_t)n + o], w = sw[i];
    Q[i] = phi * snx[i] * w;  // electric.cpp:325-328
    Q[(size_t)n + i] = phi * sny[i] * w;
    Q[2 * (size_t)n + i] = phi * snz[i] * w;
    Q[3 * (size_t)n + i] = psi * w;
}

__global__ void k_combine(int n, const int *__restrict__ perm, const double *__restrict__ snx,
                          const double *__restrict__ sny, const double *__restrict__ snz, const double *__restrict__ F,
                          const double *__restrict__ x, double *__restrict__ y, double c1, double c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    const double f0 = F[i], fx = F[(size_t)n + i], fy = F[2 * (size_t)n + i], fz = F[3 * (size_t)n + i];
    y[o] = x[o] + f0 / c1;  // electric.cpp:382-383
    y[(size_t)n + o] = x[(size_t)n + o] + (snx[i] * fx + sny[i] * fy + snz[i] * fz) / c2;
}

// the group forms: 256 points of one member per block; x and y of member m
// are picked from the kernel argument with an unrolled select (m is
// wave-uniform), which keeps the argument array out of private memory
template <typename T>
__device__ __forceinline__ T group_pick(T const (&a)[kMaxGroup], int m) {
    T r = a[0];
#pragma unroll
    for (int i = 1; i < kMaxGroup; ++i)
        if (i == m) r = a[i];
    return r;
}
__global__ __launch_bounds__(256) void k_charges_g(GroupLaunch g, GroupXY xy) {
    int bid;
    const int m = group_member(g, blockIdx.x, bid);
    const MemberView &v = g.mv[m];
    const int n = uni(v.n);
    const int i = bid * 256 + threadIdx.x;
    if (i >= n) return;
    const double *x = group_pick(xy.x, m);
    const int o = v.perm[i];
    const double phi = x[o], psi = x[(size_t)n + o], w = v.sw[i];
    double *Q = v.Q;
    Q[i] = phi * v.snx[i] * w;
    Q[(size_t)n + i] = phi * v.sny[i] * w;
    Q[2 * (size_t)n + i] = phi * v.snz[i] * w;
    Q[3 * (size_t)n + i] = psi * w;
}
__global__ __launch_bounds__(256) void k_combine_g(GroupLaunch g, GroupXY xy) {
    int bid;
    const int m = group_member(g, blockIdx.x, bid);
    const MemberView &v = g.mv[m];
    const int n = uni(v.n);
    const int i = bid * 256 + threadIdx.x;
    if (i >= n) return;
    const double *x = group_pick(xy.x, m);
    double *y = group_pick(xy.y, m);
    const double *F = v.F;
    const int o = v.perm[i];
    const double f0 = F[i], fx = F[(size_t)n + i], fy = F[2 * (size_t)n + i], fz = F[3 * (size_t)n + i];
    y[o] = x[o] + f0 / v.c1;
    y[(size_t)n + o] = x[(size_t)n + o] + (v.snx[i] * fx + v.sny[i] * fy + v.snz[i] * fz) / v.c2;
}
// L = 0 for every member (256 doubles per block)
__global__ __launch_bounds__(256) void k_zero_l_g(GroupLaunch g, int r4) {
    int bid;
    const MemberView &v = g.mv[group_member(g, blockIdx.x, bid)];
    const size_t i = (size_t)bid * 256 + threadIdx.x;
    if (i < (size_t)uni(v.nnodes) * r4) v.L[i] = 0.0;
}

// one input field from a charge vector, the others zero
__global__ void k_single_charge(int n, const int *__restrict__ perm, const double *__restrict__ q, int field,
                                double *__restrict__ Q) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = q[perm[i]];
#pragma unroll
    for (int f = 0; f < 4; ++f) Q[(size_t)f * n + i] = f == field ? v : 0.0;
}

__global__ void k_scatter_field(int n, const int *__restrict__ perm, const double *__restrict__ F, int field,
                                double sign, double *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[perm[i]] = sign * F[(size_t)field * n + i];
}

// share of the near-field work units (the costliest, they come first) that
// runs beside the upward pass of the FP16 mixed matvec
constexpr double kNearEarly = 0.15;

template <int P>
int fmm_forward_p(ibim_plan *pl, const TargetSet &ts, double *F, int nt, bool mixed) {
    constexpr int kP2mBS = P * P * P <= 128 ? 128 : 256;
    constexpr int R = P * P * P;
    constexpr int BS = R <= 64 ? 64 : (R <= 128 ? 128 : 256);
    hipStream_t st = pl->stream;
    // FP16 mixed form on the plan's own points: the near field (VALU bound)
    // runs on a second stream, its first (costliest) units beside the upward
    // pass and the rest beside the M2L reduction and L2L (memory and latency
    // bound): forked after the charges and after the M2L GEMMs, which take
    // every CU's LDS and registers, and joined before L2P
    const bool fork = mixed && P <= 6 && pl->prm.precision == IBIM_PREC_MIXED_F16 && ts.vpairs == &pl->vp &&
                      pl->f16_ready && ts.nleaves > 0 && ts.p2p;
    if (fork && !pl->stream2) {
        // all three or none, so that a failed call leaves the plan as it was
        hipStream_t s2 = nullptr;
        hipEvent_t ef = nullptr, ej = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ef, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ej, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (ej) (void)hipEventDestroy(ej);
            if (ef) (void)hipEventDestroy(ef);
            if (s2) (void)hipStreamDestroy(s2);
            set_error(std::string("second stream: ") + hipGetErrorString(e));
            return IBIM_ERR_HIP;
        }
        pl->stream2 = s2;
        pl->ev_fork = ef;
        pl->ev_join = ej;
    }
    const int ucut = fork ? (int)(ts.p2p->nunits * kNearEarly) : 0;
    // work units [u0, u1) of the near field on stream s
    auto near_field = [&](hipStream_t s, int u0, int u1) {
        const P2PWork &w = *ts.p2p;
        if (u1 <= u0) return;
        auto launch = [&](auto kern) {
            kern<<<u1 - u0, 64, 0, s>>>(w.ua + u0, w.ub + u0, ts.x, ts.y, ts.z, pl->sx, pl->sy, pl->sz, pl->Q, pl->n,
                                        ts.ranges, pl->kc, F, nt, w.part, GroupLaunch{});
        };
        const bool poly = pl->kc.mode == IBIM_KERNEL_POLY;
        if (mixed) launch(poly ? k_p2p_unit<true, true> : k_p2p_unit<true, false>);
        else launch(poly ? k_p2p_unit<false, true> : k_p2p_unit<false, false>);
    };
    const bool own = ts.p2p == &pl->p2p;  // the plan's own points (pair counts known)
    if (fork && ucut > 0) {
        IBIM_HIP_CHECK(hipEventRecord(pl->ev_fork, st));
        IBIM_HIP_CHECK(hipStreamWaitEvent(pl->stream2, pl->ev_fork, 0));
        StageScope sc(pl, kStP2PMixed, pl->stream2);
        near_field(pl->stream2, 0, ucut);
        IBIM_LAUNCH_CHECK();
        sc.add(near_field_flops(pl, 0, ucut), 0, 0);
    }
    // upward pass
    {
    StageScope sc(pl, kStP2M);
    if (pl->n_leaves_ne > 0)
        k_p2m<P, kP2mBS><<<pl->n_leaves_ne, kP2mBS, 0, st>>>(pl->leaves_ne, pl->sx, pl->sy, pl->sz, pl->Q, pl->n, pl->sbeg,
                                                   pl->send, pl->ctr, pl->level, pl->d_lvl_rad, pl->tab, pl->M);
    IBIM_LAUNCH_CHECK();
    }
    {
    StageScope sc(pl, kStM2M);
    for (int l = pl->nlevels - 2; l >= 0; --l) {
        const int cnt = pl->lvl_off[l + 1] - pl->lvl_off[l];
        // all children per pass where parents are few (latency), two where
        // they fill the chip (LDS per block sets the occupancy)
        constexpr int R4 = 4 * P * P * P;
        constexpr int GS = R4 <= 512 ? 8 : (R4 <= 1024 ? 4 : 2);  // LDS <= 64 KB
        if (cnt < 1024)
            k_m2m<P, GS><<<cnt, 256, 0, st>>>(pl->lvl_off[l], pl->flags, pl->first_child, pl->tab, pl->M);
        else
            k_m2m<P, 2><<<cnt, 256, 0, st>>>(pl->lvl_off[l], pl->flags, pl->first_child, pl->tab, pl->M);
    }
    IBIM_LAUNCH_CHECK();
    }
    // downward pass
    IBIM_HIP_CHECK(hipMemsetAsync(pl->L, 0, (size_t)pl->nnodes * 4 * R * sizeof(double), st));
    {
    StageScope sc(pl, ts.vpairs ? kStM2LX : kStM2L);
    // X list (cross-level) as class GEMMs for the plan's own points, else
    // directly; V list as GEMMs when the pairs are grouped
    const int lists = ts.vpairs ? 2 : 3;
    if (ts.xgemm) {
        int rc = m2l_gemm_x(pl, mixed);
        if (rc) return rc;
        if (!(mixed && pl->prm.precision == IBIM_PREC_MIXED_F16 && pl->f16_ready && pl->x16))
            sc.add(0, 32.0 * R * R * (double)pl->xp.npairs, 0);
    } else if (ts.n_m2l_work > 0) {
        auto launch = [&](auto kern) {
            kern<<<ts.n_m2l_work, BS, 0, st>>>(lists, ts.m2l_work, pl->loff[1], pl->litem[1], pl->loff[3],
                                               pl->litem[3], pl->flags, pl->ctr, pl->level, pl->d_lvl_rad, pl->tab,
                                               pl->kc, pl->M, pl->L);
        };
        const bool poly = pl->kc.mode == IBIM_KERNEL_POLY;
        if (mixed && ts.vpairs) launch(poly ? k_m2l_direct<P, BS, float, true> : k_m2l_direct<P, BS, float, false>);
        else launch(poly ? k_m2l_direct<P, BS, double, true> : k_m2l_direct<P, BS, double, false>);
    }
    IBIM_LAUNCH_CHECK();
    }
    if (ts.vpairs) {
        StageScope sc(pl, mixed ? kStM2LMixed : kStM2L);
        pl->defer_reduce = fork;
        int rc = m2l_gemm(pl, *ts.vpairs, mixed);
        pl->defer_reduce = false;
        if (rc) return rc;
        // (the FP16 stage also evaluates the X list)
        const bool x_here = mixed && pl->prm.precision == IBIM_PREC_MIXED_F16 && pl->f16_ready && pl->x16;
        sc.add(0, 32.0 * R * R * ((double)ts.vpairs->npairs + (x_here ? pl->xp.npairs : 0)), 0);
    }
    if (fork) {
        IBIM_HIP_CHECK(hipEventRecord(pl->ev_fork, st));
        IBIM_HIP_CHECK(hipStreamWaitEvent(pl->stream2, pl->ev_fork, 0));
        {
        StageScope sc(pl, kStP2PMixed, pl->stream2);
        near_field(pl->stream2, ucut, ts.p2p->nunits);
        IBIM_LAUNCH_CHECK();
        sc.add(near_field_flops(pl, ucut, ts.p2p->nunits), 0, 0);
        }
        IBIM_HIP_CHECK(hipEventRecord(pl->ev_join, pl->stream2));
        int rc = m2l_reduce(pl);
        if (rc) return rc;
    }
    {
    StageScope sc(pl, kStL2L);
    for (int l = 1; l < pl->nlevels; ++l) {
        const int cnt = pl->lvl_off[l + 1] - pl->lvl_off[l];
        k_l2l<P><<<cnt, 256, 0, st>>>(pl->lvl_off[l], ts.need_local, pl->parent, pl->slot, pl->tab, pl->L);
    }
    IBIM_LAUNCH_CHECK();
    }
    if (fork) IBIM_HIP_CHECK(hipStreamWaitEvent(st, pl->ev_join, 0));
    {
    StageScope sc(pl, fork ? kStL2L : mixed ? kStP2PMixed : kStP2P);
    if (ts.nleaves > 0 && ts.p2p) {
        const P2PWork &w = *ts.p2p;
        if (!fork) {
            near_field(st, 0, w.nunits);
            IBIM_LAUNCH_CHECK();
            if (own) sc.add(near_field_flops(pl, 0, w.nunits), 0, 0);
        }
        k_l2p<P><<<ts.nleaves, 64, 0, st>>>(ts.leaves, ts.tbeg, ts.tend, ts.x, ts.y, ts.z, pl->ctr, pl->level,
                                            pl->d_lvl_rad, pl->tab, pl->L, w.npart ? w.lsplit : nullptr, w.part,
                                            F, nt);
    }
    else if (ts.nleaves > 0)
        k_p2p_l2p<P><<<ts.nleaves, 64, 0, st>>>(ts.leaves, ts.tbeg, ts.tend, ts.x, ts.y, ts.z, pl->sx, pl->sy,
                                                pl->sz, pl->Q, pl->n, pl->loff[0], pl->litem[0], pl->loff[2],
                                                pl->litem[2], pl->flags, pl->sbeg, pl->send, pl->ctr, pl->level,
                                                pl->d_lvl_rad, pl->tab, pl->kc, pl->L, F, nt);
    IBIM_LAUNCH_CHECK();
    }
    return IBIM_OK;
}

}  // namespace

// merged near-field source ranges of every target leaf of a target set
int build_p2p_ranges(ibim_plan *pl, TargetSet &ts, P2PWork &work, std::vector<void *> *tmp_owner) {
    ts.roff = nullptr;
    ts.ranges = nullptr;
    ts.p2p = nullptr;
    work = P2PWork{};
    if (ts.nleaves == 0) return IBIM_OK;
    hipStream_t st = pl->stream;
    const int nl = ts.nleaves;
    auto alloc = [&](auto **p, size_t n) -> int {
        if (tmp_owner) {
            hipError_t e = dev_malloc((void **)p, (n ? n : 1) * sizeof(**p));
            if (e != hipSuccess) {
                set_error("hipMalloc (near-field work)");
                return IBIM_ERR_NOMEM;
            }
            tmp_owner->push_back((void *)*p);
            return IBIM_OK;
        }
        return dalloc(pl, p, n);
    };
    // build-time scratch, returned to the plan's pool on return
    ScratchScope scope(pl);
    auto salloc = [&](auto **p, size_t n) -> int {
        hipError_t e = smalloc(pl, p, (n ? n : 1) * sizeof(**p));
        if (e != hipSuccess) {
            set_error("hipMalloc (near-field work scratch)");
            return IBIM_ERR_NOMEM;
        }
        return IBIM_OK;
    };
    int rc;
    int *cnt = nullptr, *roff = nullptr;
    int2 *ranges = nullptr;
    if ((rc = salloc(&cnt, (size_t)nl + 1))) return rc;
    IBIM_HIP_CHECK(hipMemsetAsync(cnt, 0, ((size_t)nl + 1) * sizeof(int), st));
    k_p2p_ranges<0><<<nl, 64, 0, st>>>(nl, ts.leaves, pl->loff[0], pl->litem[0], pl->loff[2], pl->litem[2],
                                             pl->flags, pl->sbeg, pl->send, cnt, nullptr, nullptr);
    IBIM_LAUNCH_CHECK();
    std::vector<int> hc(nl);
    IBIM_HIP_CHECK(hipMemcpyAsync(hc.data(), cnt, (size_t)nl * sizeof(int), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int> ho(nl + 1, 0);
    for (int i = 0; i < nl; ++i) {
        if (hc[i] < 0) return IBIM_OK;  // more than 512 near-field boxes: keep the per-box kernel
        ho[i + 1] = ho[i] + hc[i];
    }
    if ((rc = alloc(&roff, (size_t)nl + 1)) || (rc = alloc(&ranges, (size_t)std::max(1, ho[nl])))) return rc;
    IBIM_HIP_CHECK(hipMemcpyAsync(roff, ho.data(), ho.size() * sizeof(int), hipMemcpyHostToDevice, st));
    k_p2p_ranges<1><<<nl, 64, 0, st>>>(nl, ts.leaves, pl->loff[0], pl->litem[0], pl->loff[2], pl->litem[2],
                                             pl->flags, pl->sbeg, pl->send, nullptr, roff, ranges);
    IBIM_LAUNCH_CHECK();

    // work units: count, scan, fill, order by cost (descending)
    int *ucnt = nullptr, *pcnt = nullptr, *uoff = nullptr, *poff = nullptr;
    if ((rc = salloc(&ucnt, (size_t)nl + 1)) || (rc = salloc(&pcnt, (size_t)nl + 1)) ||
        (rc = salloc(&uoff, (size_t)nl + 1)) || (rc = salloc(&poff, (size_t)nl + 1)))
        return rc;
    IBIM_HIP_CHECK(hipMemsetAsync(ucnt + nl, 0, sizeof(int), st));
    IBIM_HIP_CHECK(hipMemsetAsync(pcnt + nl, 0, sizeof(int), st));
    k_p2p_count<<<nblk(nl, 64), 64, 0, st>>>(nl, ts.leaves, ts.tbeg, ts.tend, roff, ranges, ucnt, pcnt);
    IBIM_LAUNCH_CHECK();
    size_t bytes = 0, b2 = 0;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, ucnt, uoff, nl + 1, st));
    void *tmp = nullptr;
    if ((rc = salloc((char **)&tmp, bytes))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, ucnt, uoff, nl + 1, st));
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, pcnt, poff, nl + 1, st));
    int tot[2];
    IBIM_HIP_CHECK(hipMemcpyAsync(&tot[0], uoff + nl, sizeof(int), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipMemcpyAsync(&tot[1], poff + nl, sizeof(int), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    const int nu = tot[0];
    int4 *ua = nullptr, *ub = nullptr, *sa = nullptr, *sb = nullptr;
    int *cost = nullptr, *idx = nullptr, *cost2 = nullptr, *idx2 = nullptr;
    if ((rc = salloc(&ua, nu)) || (rc = salloc(&ub, nu)) || (rc = salloc(&cost, nu)) || (rc = salloc(&idx, nu)) ||
        (rc = salloc(&cost2, nu)) || (rc = salloc(&idx2, nu)) || (rc = alloc(&sa, nu)) || (rc = alloc(&sb, nu)) ||
        (rc = alloc(&work.lsplit, nl)) || (rc = alloc(&work.part, std::max<size_t>(1, (size_t)tot[1] * 256))))
        return rc;
    k_p2p_fill<<<nblk(nl, 64), 64, 0, st>>>(nl, ts.leaves, ts.tbeg, ts.tend, roff, ranges, uoff, poff, ua, ub, cost,
                                           idx, work.lsplit);
    IBIM_LAUNCH_CHECK();
    IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b2, cost, cost2, idx, idx2, nu, 0, 31, st));
    void *tmp2 = nullptr;
    if ((rc = salloc((char **)&tmp2, b2))) return rc;
    IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(tmp2, b2, cost, cost2, idx, idx2, nu, 0, 31, st));
    k_p2p_order<<<nblk(nu, 256), 256, 0, st>>>(nu, idx2, ua, ub, sa, sb);
    IBIM_LAUNCH_CHECK();
    if (!tmp_owner && nu > 0) {  // the plan's own points: pair counts per unit prefix, on the host
        long long *pu = nullptr, *ps = nullptr;
        if ((rc = salloc(&pu, nu)) || (rc = salloc(&ps, nu))) return rc;
        k_unit_pairs<<<nblk(nu, 256), 256, 0, st>>>(nu, sa, sb, pu);
        IBIM_LAUNCH_CHECK();
        size_t b3 = 0;
        IBIM_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, b3, pu, ps, nu, st));
        void *tmp3 = nullptr;
        if ((rc = salloc((char **)&tmp3, b3))) return rc;
        IBIM_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(tmp3, b3, pu, ps, nu, st));
        work.pairs_prefix.assign((size_t)nu + 1, 0);
        IBIM_HIP_CHECK(hipMemcpyAsync(work.pairs_prefix.data() + 1, ps, (size_t)nu * sizeof(long long),
                                      hipMemcpyDeviceToHost, st));
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    work.nunits = nu;
    work.ua = sa;
    work.ub = sb;
    work.npart = (size_t)tot[1];
    ts.roff = roff;
    ts.ranges = ranges;
    ts.p2p = &work;
    return IBIM_OK;
}

double near_field_flops(const ibim_plan *pl, int u0, int u1) {
    const std::vector<int64_t> &pp = pl->p2p.pairs_prefix;
    if (pp.empty() || u1 <= u0) return 0.0;
    return 60.0 * (double)(pp[u1] - pp[u0]);
}

StageWork matvec_work(const ibim_plan *pl, bool mixed) {
    (void)mixed;
    StageWork w;
    w.valu = near_field_flops(pl, 0, pl->p2p.nunits);
    w.mfma = 32.0 * (double)pl->r * pl->r * ((double)pl->vp.npairs + pl->xp.npairs);
    // points in and out (xyz, n, w, phi, psi, two outputs, four charges) and
    // four multipole and four local vectors per non-empty box
    w.bytes = 8.0 * (14.0 * pl->n + 8.0 * pl->r * (pl->n_m2l_work + 1.0));
    return w;
}

int build_chebyshev(ibim_plan *pl) {
    const int p = pl->p;
    ChebTab &t = pl->tab;
    t = ChebTab{};
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < p; ++i) t.xbar[i] = -cos((i + 0.5) * pi / p);  // bbfmm.h:276
    for (int m = 0; m < p; ++m) {                                       // bbfmm.h:280-303
        t.Tbar[m][0] = 1.0;
        if (p > 1) t.Tbar[m][1] = t.xbar[m];
        for (int k = 2; k < p; ++k) t.Tbar[m][k] = 2.0 * t.xbar[m] * t.Tbar[m][k - 1] - t.Tbar[m][k - 2];
    }
    t.invp = 1.0 / p;
    for (int b = 0; b < 2; ++b)  // bbfmm.h:305-326
        for (int i = 0; i < p; ++i) {
            const double u = 0.5 * t.xbar[i] + (b ? 0.5 : -0.5);
            double T[kMaxP];
            T[0] = 1.0;
            if (p > 1) T[1] = u;
            for (int k = 2; k < p; ++k) T[k] = 2.0 * u * T[k - 1] - T[k - 2];
            for (int l = 0; l < p; ++l) {
                double acc = 0.0;
                for (int k = 0; k < p; ++k) acc += T[k] * t.Tbar[l][k];
                t.tr[b][i][l] = (2.0 * acc - 1.0) * t.invp;
            }
        }
    const ibim_params &q = pl->prm;
    PBConst &k = pl->kc;
    k.kappa = q.kappa;
    k.eps = q.eps_out / q.eps_in;
    k.inv_eps = q.eps_in / q.eps_out;
    k.vr = q.vacant_radius;
    k.inv4pi = 1.0 / (4.0 * pi);
    k.mode = q.kernel_mode;
    const double vr = q.vacant_radius, kap = q.kappa;  // electric.cpp:94-97
    k.k2reg = (kap == 0.0) ? 0.0 : (exp(-kap * vr) - 1.0 + kap * vr) / kap / 2.0 / pi / vr / vr;
    return IBIM_OK;
}

int fmm_forward(ibim_plan *pl, const TargetSet &ts, double *F, int nt, bool mixed) {
    switch (pl->p) {
    case 2: return fmm_forward_p<2>(pl, ts, F, nt, mixed);
    case 3: return fmm_forward_p<3>(pl, ts, F, nt, mixed);
    case 4: return fmm_forward_p<4>(pl, ts, F, nt, mixed);
    case 5: return fmm_forward_p<5>(pl, ts, F, nt, mixed);
    case 6: return fmm_forward_p<6>(pl, ts, F, nt, mixed);
    case 7: return fmm_forward_p<7>(pl, ts, F, nt, mixed);
    case 8: return fmm_forward_p<8>(pl, ts, F, nt, mixed);
    }
    set_error("unsupported Chebyshev order");
    return IBIM_ERR_ARG;
}

int form_charges(ibim_plan *pl, const double *d_x) {
    k_charges<<<nblk(pl->n), 256, 0, pl->stream>>>(pl->n, pl->perm, pl->snx, pl->sny, pl->snz, pl->sw, d_x, pl->Q);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

int combine(ibim_plan *pl, const double *d_x, double *d_y) {
    const double dE = pl->prm.eps_out, dI = pl->prm.eps_in;
    k_combine<<<nblk(pl->n), 256, 0, pl->stream>>>(pl->n, pl->perm, pl->snx, pl->sny, pl->snz, pl->F, d_x, d_y,
                                                   0.5 * (1 + dE / dI), 0.5 * (1 + dI / dE));
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

TargetSet source_targets(ibim_plan *pl) {
    TargetSet ts;
    ts.n = pl->n;
    ts.x = pl->sx;
    ts.y = pl->sy;
    ts.z = pl->sz;
    ts.tbeg = pl->sbeg;
    ts.tend = pl->send;
    ts.leaves = pl->leaves_ne;
    ts.nleaves = pl->n_leaves_ne;
    ts.m2l_work = pl->m2l_work;
    ts.n_m2l_work = pl->n_m2l_work;
    ts.need_local = pl->need_local;
    ts.vpairs = pl->ops ? &pl->vp : nullptr;
    ts.xgemm = ts.vpairs && pl->xp.npairs > 0 && pl->xops;
    ts.roff = pl->p2p_roff;
    ts.ranges = pl->p2p_ranges;
    ts.p2p = pl->p2p.ua ? &pl->p2p : nullptr;
    return ts;
}

// single kernel FMM: which charge field feeds it, which output field holds
// it and with which sign (see pair_accum for the fused combination)
static const int kFin[16] = {0, 1, 2, 3, 0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 3, 3};
static const int kFout[16] = {0, 0, 0, 0, 1, 2, 3, 2, 3, 1, 2, 3, 1, 1, 2, 3};
static const double kSgn[16] = {1, 1, 1, -1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1};

int single_charge_device(ibim_plan *pl, int kid, const double *d_q) {
    k_single_charge<<<nblk(pl->n), 256, 0, pl->stream>>>(pl->n, pl->perm, d_q, kFin[kid], pl->Q);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

int scatter_field_device(ibim_plan *pl, int kid, const int *perm, const double *F, int nt, double *d_out) {
    k_scatter_field<<<nblk(nt), 256, 0, pl->stream>>>(nt, perm, F, kFout[kid], kSgn[kid], d_out);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

int kernel_fmm_device(ibim_plan *pl, int kid, const double *d_q, double *d_out) {
    const int *fin = kFin, *fout = kFout;
    const double *sgn = kSgn;
    if (kid < 0 || kid > 15) {
        set_error("kernel_id out of range");
        return IBIM_ERR_ARG;
    }
    hipStream_t st = pl->stream;
    k_single_charge<<<nblk(pl->n), 256, 0, st>>>(pl->n, pl->perm, d_q, fin[kid], pl->Q);
    IBIM_LAUNCH_CHECK();
    TargetSet ts = source_targets(pl);
    int rc = fmm_forward(pl, ts, pl->F, pl->n);
    if (rc) return rc;
    k_scatter_field<<<nblk(pl->n), 256, 0, st>>>(pl->n, pl->perm, pl->F, fout[kid], sgn[kid], d_out);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

// ------------------------------------------------------- fused groups ---
namespace {

// the device view of a plan's matvec buffers; vops / xops are the class
// operators of the form the group runs (FP64 or FP32)
MemberView member_view(const ibim_plan *pl, int form) {
    MemberView v{};
    v.sx = pl->sx;
    v.sy = pl->sy;
    v.sz = pl->sz;
    v.snx = pl->snx;
    v.sny = pl->sny;
    v.snz = pl->snz;
    v.sw = pl->sw;
    v.perm = pl->perm;
    v.Q = pl->Q;
    v.F = pl->F;
    v.n = pl->n;
    v.nnodes = pl->nnodes;
    v.n_leaves_ne = pl->n_leaves_ne;
    v.n_m2l_work = pl->n_m2l_work;
    v.leaves_ne = pl->leaves_ne;
    v.sbeg = pl->sbeg;
    v.send = pl->send;
    v.level = pl->level;
    v.flags = pl->flags;
    v.first_child = pl->first_child;
    v.parent = pl->parent;
    v.slot = pl->slot;
    v.m2l_work = pl->m2l_work;
    v.ctr = pl->ctr;
    v.lrad = pl->d_lvl_rad;
    v.need_local = pl->need_local;
    v.M = pl->M;
    v.L = pl->L;
    v.kc = pl->kc;
    const double dE = pl->prm.eps_out, dI = pl->prm.eps_in;
    v.c1 = 0.5 * (1 + dE / dI);  // as combine()
    v.c2 = 0.5 * (1 + dI / dE);
    v.ua = pl->p2p.ua;
    v.ub = pl->p2p.ub;
    v.ranges = pl->p2p_ranges;
    v.lsplit = pl->p2p.npart ? pl->p2p.lsplit : nullptr;
    v.part = pl->p2p.part;
    const VPairs &vp = pl->vp, &xp = pl->xp;
    v.vblk = vp.blk_cls;
    v.vblk_j0 = vp.blk_j0;
    v.vkey_off = vp.cls_off_d;
    v.vA = vp.A;
    v.vB = vp.B;
    v.xblk = xp.blk_cls;
    v.xblk_j0 = xp.blk_j0;
    v.xkey_off = xp.cls_off_d;
    v.xA = xp.A;
    v.xB = xp.B;
    v.vops = form == IBIM_PREC_MIXED ? (const void *)pl->ops32 : (const void *)pl->ops;
    v.xops = form == IBIM_PREC_MIXED ? (const void *)pl->xops32 : (const void *)pl->xops;
    v.gtab = pl->gtab;
    v.stab = pl->stab;
    v.MRP = pl->MRP;
    v.KRP = pl->KRP;
    v.hvblk = vp.hblk_cls;
    v.hvblk_j0 = vp.hblk_j0;
    v.keyslot = pl->keyslot;
    v.hxblk = xp.hblk_cls;
    v.hxblk_j0 = xp.hblk_j0;
    v.hxkey_off = xp.hcls_off_d;
    v.xiota = pl->xiota;
    v.xkeyslot = pl->xkeyslot;
    v.xmask = xp.mask;
    v.ops16 = pl->ops16;
    v.esig = pl->esig;
    v.Mh = pl->Mh;
    v.Bx = pl->Bx;
    v.ebox = pl->ebox;
    v.ebx = pl->ebx;
    v.Yv = pl->Yv;
    v.y_rows = pl->y_rows;
    v.vnpairs = vp.npairs;
    v.xnpairs = xp.npairs;
    v.ytptr = pl->ytptr;
    v.ytlist = pl->ytlist;
    return v;
}

// why a plan cannot run in a group of the given form ("" if it can)
std::string group_misfit(const ibim_plan *pl, const ibim_plan *p0, int form) {
    if (!pl) return "plan is NULL";
    if (pl->device != p0->device || pl->p != p0->p || pl->kc.mode != p0->kc.mode)
        return "plans of one group share the device, the Chebyshev order and the kernel family";
    if (pl->n < 1 || !pl->p2p.ua || pl->n_leaves_ne < 1) return "plan without near-field work units";
    if (!pl->ops || (pl->xp.npairs > 0 && !pl->xops)) return "plan without class operators for its V and X lists";
    if (pl->p > 6) return "groups run p <= 6";
    if (form == IBIM_PREC_MIXED_F16) {
        if (pl->prm.precision != IBIM_PREC_MIXED_F16 || !pl->f16_ready || pl->p > 5 ||
            (pl->xp.npairs > 0 && !pl->x16))
            return "FP16-mixed groups need plans with their FP16 M2L stage built, p <= 5";
    } else if (form == IBIM_PREC_MIXED) {
        if (pl->prm.precision != IBIM_PREC_MIXED || !pl->ops32 || (pl->xp.npairs > 0 && !pl->xops32))
            return "FP32-mixed groups need plans with FP32 operators";
    }
    return "";
}

// The fused matvec, in the stage order of fmm_forward_p
template <int P>
int group_matvec_p(MatvecGroup *g, const double *const *x, double *const *y) {
    constexpr int R4 = 4 * P * P * P;
    constexpr int kP2mBS = P * P * P <= 128 ? 128 : 256;
    hipStream_t st = g->st;
    const ChebTab tab = g->pl[0]->tab;
    const bool f16 = g->form == IBIM_PREC_MIXED_F16;
    const bool mixed = g->form != IBIM_PREC_FP64;
    auto gl = [&](const MatvecGroup::Launch &L) {
        return GroupLaunch{g->d_mv, g->d_tab + L.off, L.aux >= 0 ? g->d_tab + L.aux : nullptr, g->k};
    };
    GroupXY xy{};
    for (int m = 0; m < g->k; ++m) {
        xy.x[m] = x[m];
        xy.y[m] = y[m];
    }
    auto near = [&](const MatvecGroup::Launch &L, hipStream_t s) {
        if (L.blocks == 0) return;
        auto nf = [&](auto kern) {
            kern<<<L.blocks, 64, 0, s>>>(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                         nullptr, 0, nullptr, PBConst{}, nullptr, 0, nullptr, gl(L));
        };
        if (mixed) {
            if (g->poly) nf(k_p2p_unit<true, true, true>);
            else nf(k_p2p_unit<true, false, true>);
        } else {
            if (g->poly) nf(k_p2p_unit<false, true, true>);
            else nf(k_p2p_unit<false, false, true>);
        }
    };
    StageScope sm(g->timer, mixed ? kStMatvecMixed : kStMatvec, st);
    sm.add(g->work.valu, g->work.mfma, g->work.bytes);
    if (g->charges.blocks) k_charges_g<<<g->charges.blocks, 256, 0, st>>>(gl(g->charges), xy);
    if (f16 && g->near_early.blocks) {
        IBIM_HIP_CHECK(hipEventRecord(g->ev_fork, st));
        IBIM_HIP_CHECK(hipStreamWaitEvent(g->st2, g->ev_fork, 0));
        StageScope sc(g->timer, kStP2PMixed, g->st2);
        near(g->near_early, g->st2);
        sc.add(g->near_early_flops, 0, 0);
    }
    IBIM_LAUNCH_CHECK();
    if (g->p2m.blocks)
        k_p2m<P, kP2mBS, true><<<g->p2m.blocks, kP2mBS, 0, st>>>(nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                                                 nullptr, nullptr, nullptr, nullptr, nullptr, tab,
                                                                 nullptr, gl(g->p2m));
    constexpr int GS = R4 <= 512 ? 8 : (R4 <= 1024 ? 4 : 2);
    for (const MatvecGroup::Launch &L : g->m2m) {
        if (!L.blocks) continue;
        if (L.blocks < 1024) k_m2m<P, GS, true><<<L.blocks, 256, 0, st>>>(0, nullptr, nullptr, tab, nullptr, gl(L));
        else k_m2m<P, 2, true><<<L.blocks, 256, 0, st>>>(0, nullptr, nullptr, tab, nullptr, gl(L));
    }
    if (g->zero_l.blocks) k_zero_l_g<<<g->zero_l.blocks, 256, 0, st>>>(gl(g->zero_l), R4);
    IBIM_LAUNCH_CHECK();
    {
        StageScope sc(g->timer, mixed ? kStM2LMixed : kStM2L, st);
        sc.add(0, g->m2l_flops, 0);
        if (int rc = group_m2l(g)) return rc;
    }
    if (f16) {
        IBIM_HIP_CHECK(hipEventRecord(g->ev_fork, st));
        IBIM_HIP_CHECK(hipStreamWaitEvent(g->st2, g->ev_fork, 0));
        {
            StageScope sc(g->timer, kStP2PMixed, g->st2);
            near(g->near_rest, g->st2);
            sc.add(g->near_rest_flops, 0, 0);
        }
        IBIM_LAUNCH_CHECK();
        IBIM_HIP_CHECK(hipEventRecord(g->ev_join, g->st2));
        if (int rc = group_m2l_reduce(g)) return rc;
    }
    for (const MatvecGroup::Launch &L : g->l2l)
        if (L.blocks) k_l2l<P, true><<<L.blocks, 256, 0, st>>>(0, nullptr, nullptr, nullptr, tab, nullptr, gl(L));
    IBIM_LAUNCH_CHECK();
    if (f16) {
        IBIM_HIP_CHECK(hipStreamWaitEvent(st, g->ev_join, 0));
    } else {
        StageScope sc(g->timer, mixed ? kStP2PMixed : kStP2P, st);
        near(g->near_all, st);
        sc.add(g->near_early_flops + g->near_rest_flops, 0, 0);
        IBIM_LAUNCH_CHECK();
    }
    if (g->l2p.blocks)
        k_l2p<P, true><<<g->l2p.blocks, 64, 0, st>>>(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                     nullptr, nullptr, tab, nullptr, nullptr, nullptr, nullptr, 0,
                                                     gl(g->l2p));
    if (g->combine.blocks) k_combine_g<<<g->combine.blocks, 256, 0, st>>>(gl(g->combine), xy);
    IBIM_LAUNCH_CHECK();
    return IBIM_OK;
}

}  // namespace

int group_create(ibim_plan *const *plans, int k, bool mixed, MatvecGroup **out) {
    if (!out || !plans || k < 1 || k > kMaxGroup) {
        set_error("bad argument to a matvec group (1..16 plans)");
        return IBIM_ERR_ARG;
    }
    *out = nullptr;
    const ibim_plan *p0 = plans[0];
    if (!p0) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    const int form = mixed ? p0->prm.precision : IBIM_PREC_FP64;
    for (int m = 0; m < k; ++m) {
        const std::string why = group_misfit(plans[m], p0, form);
        if (!why.empty()) {
            set_error("matvec group: " + why);
            return IBIM_ERR_ARG;
        }
        for (int q = 0; q < m; ++q)
            if (plans[q] == plans[m]) {
                set_error("matvec group: a plan appears twice");
                return IBIM_ERR_ARG;
            }
    }
    MatvecGroup *g = new MatvecGroup();
    g->k = k;
    g->p = p0->p;
    g->device = p0->device;
    g->form = form;
    g->poly = p0->kc.mode == IBIM_KERNEL_POLY;
    g->pl.assign(plans, plans + k);
    const int R4 = 4 * g->p * g->p * g->p;
    const bool f16 = form == IBIM_PREC_MIXED_F16;
    std::vector<int> &tab = g->tab;
    // one launch: blocks of member m, and an optional per-member value
    auto add = [&](MatvecGroup::Launch &L, auto blocks, auto aux, bool with_aux) {
        L.off = (int)tab.size();
        int64_t acc = 0;
        tab.push_back(0);
        for (int m = 0; m < k; ++m) {
            acc += blocks(g->pl[m]);
            tab.push_back((int)std::min<int64_t>(acc, INT32_MAX));
        }
        L.blocks = (int)std::min<int64_t>(acc, INT32_MAX);
        if (with_aux) {
            L.aux = (int)tab.size();
            for (int m = 0; m < k; ++m) tab.push_back(aux(g->pl[m]));
        }
        return acc;
    };
    auto none = [](const ibim_plan *) { return 0; };
    auto ucut = [&](const ibim_plan *pl) { return f16 ? (int)(pl->p2p.nunits * kNearEarly) : 0; };
    int64_t big = 0;
    big = std::max(big, add(g->charges, [](const ibim_plan *pl) { return nblk(pl->n); }, none, false));
    g->combine = g->charges;
    add(g->p2m, [](const ibim_plan *pl) { return pl->n_leaves_ne; }, none, false);
    big = std::max(big, add(g->zero_l, [&](const ibim_plan *pl) { return nblk((int64_t)pl->nnodes * R4); }, none,
                            false));
    if (f16) {
        add(g->near_early, ucut, [](const ibim_plan *) { return 0; }, true);
        big = std::max(big, add(g->near_rest, [&](const ibim_plan *pl) { return pl->p2p.nunits - ucut(pl); }, ucut,
                                true));
        add(g->m_to_half, [](const ibim_plan *pl) { return pl->nnodes; }, none, false);
        add(g->x_reflect, [](const ibim_plan *pl) { return pl->x16 ? pl->xp.npairs : 0; }, none, false);
        big = std::max(big, add(g->v_gemm, [](const ibim_plan *pl) { return pl->vp.hnblocks; }, none, false));
        add(g->x_gemm, [](const ibim_plan *pl) { return pl->x16 ? pl->xp.hnblocks : 0; }, none, false);
        add(g->v_reduce, [](const ibim_plan *pl) { return pl->n_m2l_work; }, none, false);
    } else {
        big = std::max(big, add(g->near_all, [](const ibim_plan *pl) { return pl->p2p.nunits; },
                                [](const ibim_plan *) { return 0; }, true));
        const int nrt = p0->MRP / 64;
        big = std::max(big, add(g->v_gemm, [&](const ibim_plan *pl) { return (int64_t)pl->vp.nblocks * nrt; }, none,
                                false));
        big = std::max(big, add(g->x_gemm, [&](const ibim_plan *pl) {
            return pl->xp.npairs > 0 ? (int64_t)pl->xp.nblocks * nrt : (int64_t)0; }, none, false));
    }
    add(g->l2p, [](const ibim_plan *pl) { return pl->n_leaves_ne; }, none, false);
    int max_lev = 0;
    for (const ibim_plan *pl : g->pl) max_lev = std::max(max_lev, pl->nlevels);
    // M2M: step s is level nlevels - 2 - s of each member (finest first)
    for (int s = 0; s + 2 <= max_lev; ++s) {
        g->m2m.emplace_back();
        auto lev = [s](const ibim_plan *pl) { return pl->nlevels - 2 - s; };
        add(g->m2m.back(), [&](const ibim_plan *pl) {
            const int l = lev(pl);
            return l >= 0 ? pl->lvl_off[l + 1] - pl->lvl_off[l] : 0; }, [&](const ibim_plan *pl) {
            const int l = lev(pl);
            return l >= 0 ? pl->lvl_off[l] : 0; }, true);
    }
    // L2L: step l is level l of each member (coarsest first)
    for (int l = 1; l < max_lev; ++l) {
        g->l2l.emplace_back();
        add(g->l2l.back(), [l](const ibim_plan *pl) { return l < pl->nlevels ? pl->lvl_off[l + 1] - pl->lvl_off[l] : 0; },
            [l](const ibim_plan *pl) { return l < pl->nlevels ? pl->lvl_off[l] : 0; }, true);
    }
    if (big >= (int64_t(1) << 23)) {
        delete g;
        set_error("matvec group: too many blocks for one fused launch (groups are for small solves)");
        return IBIM_ERR_ARG;
    }
    for (const ibim_plan *pl : g->pl) {
        const StageWork w = matvec_work(pl, mixed);
        g->work.valu += w.valu;
        g->work.mfma += w.mfma;
        g->work.bytes += w.bytes;
        g->near_early_flops += near_field_flops(pl, 0, ucut(pl));
        g->near_rest_flops += near_field_flops(pl, ucut(pl), pl->p2p.nunits);
        g->m2l_flops += 32.0 * (double)pl->r * pl->r * ((double)pl->vp.npairs + pl->xp.npairs);
    }
    // device tables, on the group's own streams
    std::vector<MemberView> mv(k);
    for (int m = 0; m < k; ++m) mv[m] = member_view(g->pl[m], form);
    hipError_t e = hipSetDevice(g->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->st2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_done, hipEventDisableTiming);
    for (int m = 0; m < k && e == hipSuccess; ++m) {
        hipEvent_t ev = nullptr;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) g->ev_in.push_back(ev);
    }
    // (from the block cache: a server makes and drops groups as its solves
    // come and go, and hipFree would wait for the whole device)
    if (e == hipSuccess) e = block_alloc((void **)&g->d_mv, k * sizeof(MemberView));
    if (e == hipSuccess) e = block_alloc((void **)&g->d_tab, tab.size() * sizeof(int));
    if (e == hipSuccess) e = hipMemcpyAsync(g->d_mv, mv.data(), k * sizeof(MemberView), hipMemcpyHostToDevice, g->st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g->d_tab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, g->st);
    if (e == hipSuccess) e = hipStreamSynchronize(g->st);
    if (e != hipSuccess) {
        set_error(std::string("matvec group set-up: ") + hipGetErrorString(e));
        group_destroy(g);
        return IBIM_ERR_HIP;
    }
    // the members' plans: the stage timers of a fused launch follow the
    // first member's selection
    g->timer.on = p0->timer.on;
    g->timer.mask = p0->timer.mask;
    *out = g;
    return IBIM_OK;
}

void group_destroy(MatvecGroup *g) {
    if (!g) return;
    if (g->st) (void)hipStreamSynchronize(g->st);
    if (g->st2) (void)hipStreamSynchronize(g->st2);
    for (auto &r : g->timer.pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t ev : g->timer.pool) (void)hipEventDestroy(ev);
    if (g->d_mv) block_release(g->d_mv);
    if (g->d_tab) block_release(g->d_tab);
    for (hipEvent_t ev : {g->ev_fork, g->ev_join, g->ev_done})
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : g->ev_in) (void)hipEventDestroy(ev);
    if (g->st2) (void)hipStreamDestroy(g->st2);
    if (g->st) (void)hipStreamDestroy(g->st);
    delete g;
}

int group_matvec_ordered(MatvecGroup *g, const double *const *x, double *const *y) {
    for (int m = 0; m < g->k; ++m) {
        IBIM_HIP_CHECK(hipEventRecord(g->ev_in[m], g->pl[m]->stream));
        IBIM_HIP_CHECK(hipStreamWaitEvent(g->st, g->ev_in[m], 0));
    }
    const int rc = group_matvec(g, x, y);
    // (also after a failed launch: whatever was queued on the group's streams
    // comes before the members' later work on their buffers)
    if (rc) {
        IBIM_HIP_CHECK(hipEventRecord(g->ev_join, g->st2));
        IBIM_HIP_CHECK(hipStreamWaitEvent(g->st, g->ev_join, 0));
        IBIM_HIP_CHECK(hipEventRecord(g->ev_done, g->st));
    }
    for (int m = 0; m < g->k; ++m) IBIM_HIP_CHECK(hipStreamWaitEvent(g->pl[m]->stream, g->ev_done, 0));
    return rc;
}

int group_matvec(MatvecGroup *g, const double *const *x, double *const *y) {
    int rc = IBIM_ERR_ARG;
    switch (g->p) {
    case 2: rc = group_matvec_p<2>(g, x, y); break;
    case 3: rc = group_matvec_p<3>(g, x, y); break;
    case 4: rc = group_matvec_p<4>(g, x, y); break;
    case 5: rc = group_matvec_p<5>(g, x, y); break;
    case 6: rc = group_matvec_p<6>(g, x, y); break;
    default: set_error("matvec group: unsupported Chebyshev order");
    }
    if (rc) return rc;
    IBIM_HIP_CHECK(hipEventRecord(g->ev_done, g->st));
    return IBIM_OK;
}

}  // namespace ibim
