// ibim_tree.hip -- the reference octree and its interaction lists, built on
// the GPU once per point set.
//
// Reference: tree::populate / getCenterRadius / assignChildren
// (bbfmm.cpp:8-138) and buildTree / buildNode / findNode / isAdjacent
// (bbfmm.cpp:141-280).
//
// Box assignment is bit-exact with the reference: a point's path is decided
// by the same comparisons `coord < centre ? 0 : 1` against the same centres,
// and centres follow the same recurrence c + (bit - 1/2) r (the product is a
// power-of-two scaling, so fused and unfused evaluation agree).  Instead of
// recursive splitting, every point gets its full path key (3 bits per level,
// child slot 4z+2y+x), the keys are radix-sorted, and the boxes of each level
// are cut out of the sorted key ranges.  A second stable sort on the keys cut
// at each point's leaf level then restores the reference's order inside a
// leaf (ascending original index, from its push_back in index order).  A box is
// split when it holds more than max_points sources and is above max_level;
// all eight children are created, empty ones become empty leaves.
#include <hipcub/hipcub.hpp>

#include "ibim_internal.h"

namespace ibim {

namespace {

constexpr int kBlk = 256;

__global__ void k_bbox_partial(const double *__restrict__ xyz, int n, double *__restrict__ part) {
    __shared__ double s[6][kBlk];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            double v = xyz[3 * (size_t)i + d];
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
        }
    }
    for (int d = 0; d < 3; ++d) {
        s[d][threadIdx.x] = mn[d];
        s[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; ++d) part[6 * blockIdx.x + d] = s[d][0];
}

__global__ void k_bbox_final(const double *__restrict__ part, int g, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < g; ++b)
        for (int d = 0; d < 3; ++d) {
            v[d] = fmin(v[d], part[6 * b + d]);
            v[3 + d] = fmax(v[3 + d], part[6 * b + 3 + d]);
        }
    for (int d = 0; d < 6; ++d) out[d] = v[d];
}

// full path key of every point: one 3-bit child slot per level
__global__ void k_point_keys(const double *__restrict__ xyz, int n, double cx, double cy, double cz, double rx,
                             double ry, double rz, int maxlev, uint64_t *__restrict__ keys, int *__restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double px = xyz[3 * (size_t)i], py = xyz[3 * (size_t)i + 1], pz = xyz[3 * (size_t)i + 2];
    uint64_t key = 0;
    for (int l = 0; l < maxlev; ++l) {
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

__global__ void k_gather_geometry(const int *__restrict__ perm, int n, const double *__restrict__ xyz,
                                  const double *__restrict__ nrm, const double *__restrict__ w, double *sx,
                                  double *sy, double *sz, double *snx, double *sny, double *snz, double *sw) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    size_t o = (size_t)perm[i];
    sx[i] = xyz[3 * o];
    sy[i] = xyz[3 * o + 1];
    sz[i] = xyz[3 * o + 2];
    if (nrm) {
        snx[i] = nrm[3 * o];
        sny[i] = nrm[3 * o + 1];
        snz[i] = nrm[3 * o + 2];
    }
    if (w) sw[i] = w[o];
}

__global__ void k_split_flags(const int *__restrict__ beg, const int *__restrict__ end, int nl, int level,
                              int max_points, int max_level, int *__restrict__ flag) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    int cnt = end[j] - beg[j];
    flag[j] = (cnt > 0 && cnt > max_points && level < max_level) ? 1 : 0;
}

// digit of level (l+1) of a full key
__device__ __forceinline__ int key_digit(uint64_t key, int shift) { return (int)((key >> shift) & 7u); }

__global__ void k_make_children(const int *__restrict__ flag, const int *__restrict__ base, int nl,
                                const uint64_t *__restrict__ pkey, const int *__restrict__ pbeg,
                                const int *__restrict__ pend, const double *__restrict__ pctr, double prx,
                                double pry, double prz, const uint64_t *__restrict__ skeys, int shift,
                                uint64_t *__restrict__ ckey, int *__restrict__ cbeg, int *__restrict__ cend,
                                double *__restrict__ cctr, int *__restrict__ cpar, int *__restrict__ cslot) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl || !flag[j]) return;
    int b0 = pbeg[j], e0 = pend[j];
    int lb[9];
    lb[0] = b0;
    lb[8] = e0;
    for (int c = 1; c < 8; ++c) {  // first index in [b0,e0) whose digit >= c
        int lo = b0, hi = e0;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (key_digit(skeys[mid], shift) < c) lo = mid + 1;
            else hi = mid;
        }
        lb[c] = lo;
    }
    double pcx = pctr[3 * j], pcy = pctr[3 * j + 1], pcz = pctr[3 * j + 2];
    for (int c = 0; c < 8; ++c) {
        int k = 8 * base[j] + c;
        ckey[k] = (pkey[j] << 3) | (uint64_t)c;
        cbeg[k] = lb[c];
        cend[k] = lb[c + 1];
        // bbfmm.cpp:93-95
        cctr[3 * k] = pcx + ((double)(c & 1) - 0.5) * prx;
        cctr[3 * k + 1] = pcy + ((double)((c >> 1) & 1) - 0.5) * pry;
        cctr[3 * k + 2] = pcz + ((double)(c >> 2) - 0.5) * prz;
        cpar[k] = j;
        cslot[k] = c;
    }
}

__global__ void k_concat_level(int nl, int off, int level, int child_off, int parent_off, const int *__restrict__ flag,
                               const int *__restrict__ base, const uint64_t *__restrict__ lkey,
                               const int *__restrict__ lbeg, const int *__restrict__ lend,
                               const double *__restrict__ lctr, const int *__restrict__ lpar,
                               const int *__restrict__ lslot, uint64_t *nkey, int *nlev, int *npar, int *nfc,
                               int *nslot, int *nflags, int *nsb, int *nse, double *nctr) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    int id = off + j;
    nkey[id] = lkey[j];
    nlev[id] = level;
    npar[id] = lpar ? parent_off + lpar[j] : -1;
    nslot[id] = lslot ? lslot[j] : 0;
    int split = flag[j];
    nfc[id] = split ? child_off + 8 * base[j] : -1;
    int cnt = lend[j] - lbeg[j];
    nflags[id] = (split ? 0 : kLeaf) | (cnt == 0 ? kEmpty : 0);
    nsb[id] = lbeg[j];
    nse[id] = lend[j];
    nctr[3 * id] = lctr[3 * j];
    nctr[3 * id + 1] = lctr[3 * j + 1];
    nctr[3 * id + 2] = lctr[3 * j + 2];
}

// key of each point truncated at its leaf's level, written at the point's
// original index: a stable sort on it keeps the leaves in the same order and
// puts the points of a leaf in ascending original index (bbfmm.cpp:106-116)
__global__ void k_leaf_keys(int nnodes, const int *__restrict__ flags, const uint64_t *__restrict__ nkey,
                            const int *__restrict__ level, const int *__restrict__ sb, const int *__restrict__ se,
                            const int *__restrict__ perm, int maxlev, uint64_t *__restrict__ tkey) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nnodes || !(flags[id] & kLeaf)) return;
    const uint64_t k = nkey[id] << (3 * (maxlev - level[id]));
    for (int i = sb[id]; i < se[id]; ++i) tkey[perm[i]] = k;
}

__global__ void k_iota(int n, int *__restrict__ a) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// ------------------------------------------------------------ lists ----
struct TreeView {
    const double *ctr;
    const double *lrad;  // 3 per level
    const int *level, *parent, *fc, *flags;
};

// bbfmm.cpp:244-258
__device__ int find_node(const TreeView &t, double px, double py, double pz) {
    int id = 0;
    for (;;) {
        double cx = t.ctr[3 * id], cy = t.ctr[3 * id + 1], cz = t.ctr[3 * id + 2];
        if (fabs(cx - px) < kEps && fabs(cy - py) < kEps && fabs(cz - pz) < kEps) return id;
        if (t.flags[id] & kLeaf) return id;
        int xb = cx > px ? 0 : 1;
        int yb = cy > py ? 0 : 1;
        int zb = cz > pz ? 0 : 1;
        id = t.fc[id] + 4 * zb + 2 * yb + xb;
    }
}

// bbfmm.cpp:260-280
__device__ bool is_adjacent(const TreeView &t, int a, int b) {
    const double *ra = t.lrad + 3 * t.level[a], *rb = t.lrad + 3 * t.level[b];
    double dx = fabs(t.ctr[3 * a] - t.ctr[3 * b]);
    double dy = fabs(t.ctr[3 * a + 1] - t.ctr[3 * b + 1]);
    double dz = fabs(t.ctr[3 * a + 2] - t.ctr[3 * b + 2]);
    double rx = fabs(ra[0] + rb[0]), ry = fabs(ra[1] + rb[1]), rz = fabs(ra[2] + rb[2]);
    bool rdx = rx >= dx - kEps, rdy = ry >= dy - kEps, rdz = rz >= dz - kEps;
    bool xa = (fabs(dx - rx) < kEps) && (rdy && rdz);
    bool ya = (fabs(dy - ry) < kEps) && (rdx && rdz);
    bool za = (fabs(dz - rz) < kEps) && (rdy && rdx);
    return xa || ya || za;
}

struct ListSink {
    int *cnt;        // pass 0: counts [4][nnodes]
    const int *off;  // pass 1: offsets [4] arrays
    int *item[4];
    const int *offs[4];
    int nnodes;
};

// bbfmm.cpp:158-241; pass 0 counts, pass 1 writes
template <int PASS>
__global__ void k_lists(TreeView t, int nnodes, double minx, double miny, double minz, double maxx, double maxy,
                        double maxz, int *__restrict__ counts, const int *__restrict__ uoff,
                        const int *__restrict__ voff, const int *__restrict__ woff, const int *__restrict__ xoff,
                        int *__restrict__ uit, int *__restrict__ vit, int *__restrict__ wit, int *__restrict__ xit) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nnodes) return;
    int nu = 0, nv = 0, nw = 0, nx = 0;
    int ub = 0, vb = 0, wb = 0, xb = 0;
    if (PASS == 1) {
        ub = uoff[id];
        vb = voff[id];
        wb = woff[id];
        xb = xoff[id];
    }
    auto emit = [&](int kind, int v) {
        if (PASS == 1) {
            if (kind == 0) uit[ub + nu] = v;
            else if (kind == 1) vit[vb + nv] = v;
            else if (kind == 2) wit[wb + nw] = v;
            else xit[xb + nx] = v;
        }
        if (kind == 0) ++nu;
        else if (kind == 1) ++nv;
        else if (kind == 2) ++nw;
        else ++nx;
    };
    const bool leaf = (t.flags[id] & kLeaf) != 0;
    const int lev = t.level[id];
    const int par = t.parent[id];
    if (par != -1) {
        const double *rd = t.lrad + 3 * lev;
        double dx = rd[0], dy = rd[1], dz = rd[2];
        double xs = t.ctr[3 * par] - dx, ys = t.ctr[3 * par + 1] - dy, zs = t.ctr[3 * par + 2] - dz;
        double mcx = t.ctr[3 * id], mcy = t.ctr[3 * id + 1], mcz = t.ctr[3 * id + 2];
        int coarse[32];
        int ncoarse = 0;
        int stack[192];
        for (int xi = -2; xi < 4; ++xi)
            for (int yi = -2; yi < 4; ++yi)
                for (int zi = -2; zi < 4; ++zi) {
                    double cx = xs + (double)(2 * xi) * dx;
                    double cy = ys + (double)(2 * yi) * dy;
                    double cz = zs + (double)(2 * zi) * dz;
                    bool inside = cx <= maxx + kEps && cy <= maxy + kEps && cz <= maxz + kEps && cx >= minx - kEps &&
                                  cy >= miny - kEps && cz >= minz - kEps;
                    bool self = fabs(cx - mcx) < kEps && fabs(cy - mcy) < kEps && fabs(cz - mcz) < kEps;
                    if (!inside || self) continue;
                    int cid = find_node(t, cx, cy, cz);
                    bool adj = is_adjacent(t, id, cid);
                    int clev = t.level[cid];
                    if (clev < lev) {
                        bool seen = false;
                        for (int k = 0; k < ncoarse; ++k) seen |= (coarse[k] == cid);
                        if (!seen) {
                            if (ncoarse < 32) coarse[ncoarse++] = cid;
                            if (adj) {
                                if (t.flags[cid] & kLeaf) emit(0, cid);
                            } else {
                                emit(3, cid);
                            }
                        }
                    }
                    if (clev == lev) {
                        if (!adj) {
                            emit(1, cid);
                        } else if (leaf) {
                            int sp = 0;
                            stack[sp++] = cid;
                            while (sp > 0) {
                                int f = stack[--sp];
                                if (!is_adjacent(t, f, id)) {
                                    emit(2, f);
                                } else if (t.flags[f] & kLeaf) {
                                    emit(0, f);
                                } else if (sp + 8 <= 192) {
                                    int fc = t.fc[f];
                                    for (int c = 7; c >= 0; --c) stack[sp++] = fc + c;
                                }
                            }
                        }
                    }
                }
    }
    if (leaf) emit(0, id);
    if (PASS == 0) {
        counts[id] = nu;
        counts[nnodes + id] = nv;
        counts[2 * nnodes + id] = nw;
        counts[3 * nnodes + id] = nx;
    }
}

__global__ void k_flag_nonempty_leaf(const int *__restrict__ flags, int nnodes, unsigned char *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    out[i] = ((flags[i] & kLeaf) && !(flags[i] & kEmpty)) ? 1 : 0;
}

__global__ void k_flag_nonempty_nonroot(const int *__restrict__ flags, int nnodes, unsigned char *__restrict__ out,
                                        unsigned char *__restrict__ need) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnodes) return;
    bool ne = !(flags[i] & kEmpty);
    out[i] = (ne && i != 0) ? 1 : 0;
    need[i] = ne ? 1 : 0;
}

inline int nblk(int64_t n, int b = kBlk) { return (int)((n + b - 1) / b); }

}  // namespace

// exclusive scan of n ints into out (n+1 entries, out[n] = total)
static int scan_ints(ibim_plan *pl, const int *in, int *out, int n, int *total) {
    ScratchScope scope(pl);
    size_t bytes = 0;
    IBIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n + 1, pl->stream));
    void *tmp = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &tmp, bytes ? bytes : 1));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n + 1, pl->stream);
    if (e == hipSuccess && total) e = hipMemcpyAsync(total, out + n, sizeof(int), hipMemcpyDeviceToHost, pl->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
    if (e != hipSuccess) {
        set_error(std::string("scan: ") + hipGetErrorString(e));
        return IBIM_ERR_HIP;
    }
    return IBIM_OK;
}

static int select_flagged(ibim_plan *pl, const unsigned char *flags, int n, int **out, int *count) {
    ScratchScope scope(pl);
    int rc = dalloc(pl, out, (size_t)n);
    if (rc) return rc;
    int *d_num = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &d_num, sizeof(int)));
    size_t bytes = 0;
    hipcub::CountingInputIterator<int> it(0);
    IBIM_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flags, *out, d_num, n, pl->stream));
    void *tmp = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &tmp, bytes ? bytes : 1));
    hipError_t e = hipcub::DeviceSelect::Flagged(tmp, bytes, it, flags, *out, d_num, n, pl->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(count, d_num, sizeof(int), hipMemcpyDeviceToHost, pl->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
    if (e != hipSuccess) {
        set_error(std::string("select: ") + hipGetErrorString(e));
        return IBIM_ERR_HIP;
    }
    return IBIM_OK;
}

int build_tree(ibim_plan *pl, const double *d_xyz, const double *d_nrm, const double *d_w) {
    ScratchScope scope(pl);  // every temporary below
    const int n = pl->n;
    const int maxlev = pl->prm.max_level;
    hipStream_t st = pl->stream;
    int rc;

    // ---- root box: bounding box of the sources (bbfmm.cpp:43-65)
    int g = std::min(1024, std::max(1, nblk(n)));
    double *part = nullptr, *bb = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &part, 6 * (size_t)g * sizeof(double)));
    IBIM_HIP_CHECK(smalloc(pl, &bb, 6 * sizeof(double)));
    k_bbox_partial<<<g, kBlk, 0, st>>>(d_xyz, n, part);
    k_bbox_final<<<1, 64, 0, st>>>(part, g, bb);
    IBIM_LAUNCH_CHECK();
    double hb[6];
    IBIM_HIP_CHECK(hipMemcpyAsync(hb, bb, sizeof(hb), hipMemcpyDeviceToHost, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    for (int d = 0; d < 3; ++d) {
        if (!std::isfinite(hb[d]) || !std::isfinite(hb[3 + d])) {
            set_error("non-finite point coordinates");
            return IBIM_ERR_ARG;
        }
        pl->root_c[d] = (hb[3 + d] + hb[d]) / 2.0;
        pl->root_r[d] = (hb[3 + d] - hb[d]) / 2.0;
        if (!(pl->root_r[d] > 0.0)) {
            // the reference divides by this radius (bbfmm.h:403-405) and
            // would return NaN; refuse instead
            set_error("the points span zero extent along an axis");
            return IBIM_ERR_ARG;
        }
    }

    // ---- keys and stable sort
    uint64_t *k_in = nullptr, *k_out = nullptr;
    int *i_in = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &k_in, (size_t)n * sizeof(uint64_t)));
    IBIM_HIP_CHECK(smalloc(pl, &k_out, (size_t)n * sizeof(uint64_t)));
    IBIM_HIP_CHECK(smalloc(pl, &i_in, (size_t)n * sizeof(int)));
    if ((rc = dalloc(pl, &pl->perm, (size_t)n))) return rc;
    k_point_keys<<<nblk(n), kBlk, 0, st>>>(d_xyz, n, pl->root_c[0], pl->root_c[1], pl->root_c[2], pl->root_r[0],
                                           pl->root_r[1], pl->root_r[2], maxlev, k_in, i_in);
    IBIM_LAUNCH_CHECK();
    if (maxlev > 0) {
        size_t bytes = 0;
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k_in, k_out, i_in, pl->perm, n, 0,
                                                          3 * maxlev, st));
        void *tmp = nullptr;
        IBIM_HIP_CHECK(smalloc(pl, &tmp, bytes ? bytes : 1));
        IBIM_HIP_CHECK(
            hipcub::DeviceRadixSort::SortPairs(tmp, bytes, k_in, k_out, i_in, pl->perm, n, 0, 3 * maxlev, st));
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
    } else {
        IBIM_HIP_CHECK(hipMemcpyAsync(k_out, k_in, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(pl->perm, i_in, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, st));
    }

    // ---- sorted geometry
    if ((rc = dalloc(pl, &pl->sx, n)) || (rc = dalloc(pl, &pl->sy, n)) || (rc = dalloc(pl, &pl->sz, n)) ||
        (rc = dalloc(pl, &pl->snx, n)) || (rc = dalloc(pl, &pl->sny, n)) || (rc = dalloc(pl, &pl->snz, n)) ||
        (rc = dalloc(pl, &pl->sw, n)))
        return rc;
    k_gather_geometry<<<nblk(n), kBlk, 0, st>>>(pl->perm, n, d_xyz, d_nrm, d_w, pl->sx, pl->sy, pl->sz, pl->snx,
                                                pl->sny, pl->snz, pl->sw);
    IBIM_LAUNCH_CHECK();

    // ---- boxes, level by level
    struct Level {
        int nl = 0;
        uint64_t *key = nullptr;
        int *beg = nullptr, *end = nullptr, *par = nullptr, *slot = nullptr, *flag = nullptr, *base = nullptr;
        double *ctr = nullptr;
    };
    std::vector<Level> lv;
    auto talloc = [&](auto **p, size_t cnt) -> int {
        hipError_t e = smalloc(pl, p, (cnt ? cnt : 1) * sizeof(**p));
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc (tree build): ") + hipGetErrorString(e));
            return IBIM_ERR_NOMEM;
        }
        return IBIM_OK;
    };
    auto cleanup = [&]() {};  // scratch returns to the pool with the scope
    {
        Level r0;
        r0.nl = 1;
        if ((rc = talloc(&r0.key, 1)) || (rc = talloc(&r0.beg, 1)) || (rc = talloc(&r0.end, 1)) ||
            (rc = talloc(&r0.ctr, 3))) {
            cleanup();
            return rc;
        }
        uint64_t k0 = 0;
        int b0 = 0, e0 = n;
        IBIM_HIP_CHECK(hipMemcpyAsync(r0.key, &k0, sizeof(k0), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(r0.beg, &b0, sizeof(int), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(r0.end, &e0, sizeof(int), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(r0.ctr, pl->root_c, 3 * sizeof(double), hipMemcpyHostToDevice, st));
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
        lv.push_back(r0);
    }
    pl->lvl_rad.clear();
    double rad[3] = {pl->root_r[0], pl->root_r[1], pl->root_r[2]};
    for (int l = 0;; ++l) {
        Level &cur = lv[l];
        pl->lvl_rad.insert(pl->lvl_rad.end(), rad, rad + 3);
        if ((rc = talloc(&cur.flag, (size_t)cur.nl + 1)) || (rc = talloc(&cur.base, (size_t)cur.nl + 1))) {
            cleanup();
            return rc;
        }
        IBIM_HIP_CHECK(hipMemsetAsync(cur.flag, 0, ((size_t)cur.nl + 1) * sizeof(int), st));
        k_split_flags<<<nblk(cur.nl), kBlk, 0, st>>>(cur.beg, cur.end, cur.nl, l, pl->prm.max_points, maxlev,
                                                     cur.flag);
        IBIM_LAUNCH_CHECK();
        int nsplit = 0;
        if ((rc = scan_ints(pl, cur.flag, cur.base, cur.nl, &nsplit))) {
            cleanup();
            return rc;
        }
        if (nsplit == 0) break;
        Level nx;
        nx.nl = 8 * nsplit;
        if ((rc = talloc(&nx.key, nx.nl)) || (rc = talloc(&nx.beg, nx.nl)) || (rc = talloc(&nx.end, nx.nl)) ||
            (rc = talloc(&nx.par, nx.nl)) || (rc = talloc(&nx.slot, nx.nl)) || (rc = talloc(&nx.ctr, 3 * (size_t)nx.nl))) {
            cleanup();
            return rc;
        }
        int shift = 3 * (maxlev - (l + 1));
        k_make_children<<<nblk(cur.nl), kBlk, 0, st>>>(cur.flag, cur.base, cur.nl, cur.key, cur.beg, cur.end, cur.ctr,
                                                       rad[0], rad[1], rad[2], k_out, shift, nx.key, nx.beg, nx.end,
                                                       nx.ctr, nx.par, nx.slot);
        IBIM_LAUNCH_CHECK();
        for (int d = 0; d < 3; ++d) rad[d] = rad[d] * 0.5;
        lv.push_back(nx);
    }
    pl->nlevels = (int)lv.size();
    pl->lvl_off.assign(pl->nlevels + 1, 0);
    for (int l = 0; l < pl->nlevels; ++l) pl->lvl_off[l + 1] = pl->lvl_off[l] + lv[l].nl;
    pl->nnodes = pl->lvl_off[pl->nlevels];
    const int nn = pl->nnodes;
    if ((rc = dalloc(pl, &pl->nkey, nn)) || (rc = dalloc(pl, &pl->level, nn)) || (rc = dalloc(pl, &pl->parent, nn)) ||
        (rc = dalloc(pl, &pl->first_child, nn)) || (rc = dalloc(pl, &pl->slot, nn)) ||
        (rc = dalloc(pl, &pl->flags, nn)) || (rc = dalloc(pl, &pl->sbeg, nn)) || (rc = dalloc(pl, &pl->send, nn)) ||
        (rc = dalloc(pl, &pl->ctr, 3 * (size_t)nn)) || (rc = dalloc(pl, &pl->d_lvl_rad, 3 * (size_t)pl->nlevels))) {
        cleanup();
        return rc;
    }
    for (int l = 0; l < pl->nlevels; ++l) {
        const Level &L = lv[l];
        int child_off = l + 1 < pl->nlevels ? pl->lvl_off[l + 1] : 0;
        int parent_off = l > 0 ? pl->lvl_off[l - 1] : 0;
        k_concat_level<<<nblk(L.nl), kBlk, 0, st>>>(L.nl, pl->lvl_off[l], l, child_off, parent_off, L.flag, L.base,
                                                    L.key, L.beg, L.end, L.ctr, L.par, L.slot, pl->nkey, pl->level,
                                                    pl->parent, pl->first_child, pl->slot, pl->flags, pl->sbeg,
                                                    pl->send, pl->ctr);
        IBIM_LAUNCH_CHECK();
    }
    IBIM_HIP_CHECK(hipMemcpyAsync(pl->d_lvl_rad, pl->lvl_rad.data(), pl->lvl_rad.size() * sizeof(double),
                                  hipMemcpyHostToDevice, st));
    IBIM_HIP_CHECK(hipStreamSynchronize(st));
    cleanup();

    // ---- points of a leaf in ascending original index: second stable sort on
    // the leaf-truncated keys, starting from the original order
    if (maxlev > 0) {
        uint64_t *tk = nullptr;
        int *iota = nullptr, *perm2 = nullptr;
        IBIM_HIP_CHECK(smalloc(pl, &tk, (size_t)n * sizeof(uint64_t)));
        IBIM_HIP_CHECK(smalloc(pl, &iota, (size_t)n * sizeof(int)));
        IBIM_HIP_CHECK(smalloc(pl, &perm2, (size_t)n * sizeof(int)));
        k_leaf_keys<<<nblk(nn), kBlk, 0, st>>>(nn, pl->flags, pl->nkey, pl->level, pl->sbeg, pl->send, pl->perm,
                                               maxlev, tk);
        k_iota<<<nblk(n), kBlk, 0, st>>>(n, iota);
        IBIM_LAUNCH_CHECK();
        size_t bytes = 0;
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, tk, k_out, iota, perm2, n, 0, 3 * maxlev, st));
        void *tmp2 = nullptr;
        IBIM_HIP_CHECK(smalloc(pl, &tmp2, bytes ? bytes : 1));
        IBIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp2, bytes, tk, k_out, iota, perm2, n, 0, 3 * maxlev, st));
        IBIM_HIP_CHECK(hipMemcpyAsync(pl->perm, perm2, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, st));
        k_gather_geometry<<<nblk(n), kBlk, 0, st>>>(pl->perm, n, d_xyz, d_nrm, d_w, pl->sx, pl->sy, pl->sz, pl->snx,
                                                    pl->sny, pl->snz, pl->sw);
        IBIM_LAUNCH_CHECK();
        IBIM_HIP_CHECK(hipStreamSynchronize(st));
    }

    // ---- interaction lists (count, scan, fill)
    TreeView tv{pl->ctr, pl->d_lvl_rad, pl->level, pl->parent, pl->first_child, pl->flags};
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
        mn[d] = pl->root_c[d] - pl->root_r[d];
        mx[d] = pl->root_c[d] + pl->root_r[d];
    }
    int *counts = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &counts, 4 * ((size_t)nn + 1) * sizeof(int)));
    IBIM_HIP_CHECK(hipMemsetAsync(counts, 0, 4 * ((size_t)nn + 1) * sizeof(int), st));
    k_lists<0><<<nblk(nn, 64), 64, 0, st>>>(tv, nn, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], counts, nullptr,
                                            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    IBIM_LAUNCH_CHECK();
    // counts laid out [4][nn]; scan each kind separately (needs n+1 input entries)
    int *cnt_pad = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &cnt_pad, ((size_t)nn + 1) * sizeof(int)));
    for (int k = 0; k < 4; ++k) {
        if ((rc = dalloc(pl, &pl->loff[k], (size_t)nn + 1))) return rc;
        IBIM_HIP_CHECK(hipMemsetAsync(cnt_pad, 0, ((size_t)nn + 1) * sizeof(int), st));
        IBIM_HIP_CHECK(hipMemcpyAsync(cnt_pad, counts + (size_t)k * nn, (size_t)nn * sizeof(int),
                                      hipMemcpyDeviceToDevice, st));
        int tot = 0;
        if ((rc = scan_ints(pl, cnt_pad, pl->loff[k], nn, &tot))) return rc;
        pl->ltot[k] = tot;
        if ((rc = dalloc(pl, &pl->litem[k], (size_t)tot))) return rc;
    }
    k_lists<1><<<nblk(nn, 64), 64, 0, st>>>(tv, nn, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], nullptr, pl->loff[0],
                                            pl->loff[1], pl->loff[2], pl->loff[3], pl->litem[0], pl->litem[1],
                                            pl->litem[2], pl->litem[3]);
    IBIM_LAUNCH_CHECK();

    // ---- work lists for the points as their own targets
    unsigned char *fl = nullptr;
    IBIM_HIP_CHECK(smalloc(pl, &fl, (size_t)nn));
    if ((rc = dalloc(pl, &pl->need_local, (size_t)nn))) return rc;
    k_flag_nonempty_leaf<<<nblk(nn), kBlk, 0, st>>>(pl->flags, nn, fl);
    IBIM_LAUNCH_CHECK();
    if ((rc = select_flagged(pl, fl, nn, &pl->leaves_ne, &pl->n_leaves_ne))) return rc;
    k_flag_nonempty_nonroot<<<nblk(nn), kBlk, 0, st>>>(pl->flags, nn, fl, pl->need_local);
    IBIM_LAUNCH_CHECK();
    if ((rc = select_flagged(pl, fl, nn, &pl->m2l_work, &pl->n_m2l_work))) return rc;
    IBIM_HIP_CHECK(hipStreamSynchronize(st));

    // deepest non-empty leaf
    std::vector<int> hflags(nn);
    IBIM_HIP_CHECK(hipMemcpy(hflags.data(), pl->flags, (size_t)nn * sizeof(int), hipMemcpyDeviceToHost));
    pl->depth = 0;
    for (int l = 0; l < pl->nlevels; ++l)
        for (int i = pl->lvl_off[l]; i < pl->lvl_off[l + 1]; ++i)
            if ((hflags[i] & kLeaf) && !(hflags[i] & kEmpty)) pl->depth = l;
    return IBIM_OK;
}

}  // namespace ibim
