// ibim_capi.hip -- the C ABI of include/ibim.h over the device stages.
#include <cmath>
#include <cstring>

#include <map>
#include <mutex>
#include <unordered_map>

#include "ibim_internal.h"

namespace ibim {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

int kernel_fmm_device(ibim_plan *pl, int kid, const double *d_q, double *d_out);
int polarize_device(ibim_plan *pl, int m, const double *d_axyz, const double *d_x, double *d_out);
int polarize_noalloc(ibim_plan *pl, int m, const double *d_axyz, const double *d_x, double *d_out);
int target_fmm_device(ibim_plan *pl, int kid, int m, const double *d_axyz, const double *d_in, double *d_out);

int timer_collect(StageTimer &t) {
    for (auto &r : t.pending) {
        IBIM_HIP_CHECK(hipEventSynchronize(r.b));
        float el = 0.f;
        IBIM_HIP_CHECK(hipEventElapsedTime(&el, r.a, r.b));
        t.ms[r.stage] += el;
        t.count[r.stage] += 1;
        t.work[r.stage].valu += r.w.valu;
        t.work[r.stage].mfma += r.w.mfma;
        t.work[r.stage].bytes += r.w.bytes;
        t.pool.push_back(r.a);
        t.pool.push_back(r.b);
    }
    t.pending.clear();
    return IBIM_OK;
}

int timer_report(StageTimer &t, ibim_stage_report *out, int reset) {
    if (int rc = timer_collect(t)) return rc;
    if (out) {
        *out = ibim_stage_report{};
        for (int i = 0; i < IBIM_STAGE_COUNT; ++i) {
            out->ms[i] = t.ms[i];
            out->launches[i] = t.count[i];
            out->valu_flops[i] = t.work[i].valu;
            out->mfma_flops[i] = t.work[i].mfma;
            out->bytes[i] = t.work[i].bytes;
        }
    }
    if (reset)
        for (int i = 0; i < 16; ++i) {
            t.ms[i] = 0;
            t.count[i] = 0;
            t.work[i] = StageWork{};
        }
    return IBIM_OK;
}

// one dmapping (electric.cpp:294-388): charges, fused FMM, combine
int matvec_device(ibim_plan *pl, const double *d_x, double *d_y, bool mixed) {
    if (pl->batcher) return batch_matvec(pl->batcher, pl, d_x, d_y, mixed);
    return matvec_single(pl, d_x, d_y, mixed);
}

int matvec_single(ibim_plan *pl, const double *d_x, double *d_y, bool mixed) {
    if (mixed && pl->prm.inner_order > 0) {
        // a plan whose inner solves use a lower order builds its own mixed
        // operators on its first mixed matvec (they both return at once when built)
        int rc = pl->prm.precision == IBIM_PREC_MIXED_F16 ? build_m2l_operators_f16(pl)
                                                          : build_m2l_operators_f32(pl);
        if (rc) return rc;
    }
    StageScope sc(pl, mixed ? kStMatvecMixed : kStMatvec);
    const StageWork w = matvec_work(pl, mixed);
    sc.add(w.valu, w.mfma, w.bytes);
    int rc = form_charges(pl, d_x);
    if (rc) return rc;
    TargetSet ts = source_targets(pl);
    if ((rc = fmm_forward(pl, ts, pl->F, pl->n, mixed))) return rc;
    return combine(pl, d_x, d_y);
}

static int check_params(const ibim_params *q, int64_t n) {
    if (!q) {
        set_error("params is NULL");
        return IBIM_ERR_ARG;
    }
    if (n < 1 || n > (int64_t)0x7fffffff) {
        set_error("number of points must be in 1..2^31-1");
        return IBIM_ERR_ARG;
    }
    if (q->p < 2 || q->p > kMaxP) {
        set_error("Chebyshev order p must be in 2..8");
        return IBIM_ERR_ARG;
    }
    if (q->max_points < 1) {
        set_error("max_points must be >= 1");
        return IBIM_ERR_ARG;
    }
    if (q->max_level < 0 || q->max_level > 20) {
        set_error("max_level must be in 0..20");
        return IBIM_ERR_ARG;
    }
    if (!(q->eps_in > 0) || !(q->eps_out > 0) || !(q->kappa >= 0)) {
        set_error("eps_in, eps_out must be > 0 and kappa >= 0");
        return IBIM_ERR_ARG;
    }
    if (q->kernel_mode != IBIM_KERNEL_PB && q->kernel_mode != IBIM_KERNEL_POLY) {
        set_error("unknown kernel_mode");
        return IBIM_ERR_ARG;
    }
    if (q->kernel_mode == IBIM_KERNEL_PB && !(q->vacant_radius > 0)) {
        // the reference divides by r = 0 at the self pair when tau*h <= 0
        set_error("vacant_radius (tau*h) must be > 0 for the PB kernels");
        return IBIM_ERR_ARG;
    }
    if (q->precision != IBIM_PREC_FP64 && q->precision != IBIM_PREC_MIXED && q->precision != IBIM_PREC_MIXED_F16) {
        set_error("unknown precision");
        return IBIM_ERR_ARG;
    }
    if (q->inner_order != 0 && (q->inner_order < 2 || q->inner_order >= q->p)) {
        set_error("inner_order must be 0 or in 2..p-1");
        return IBIM_ERR_ARG;
    }
    return IBIM_OK;
}

}  // namespace ibim

using namespace ibim;

// host-array wrapper: copy in, run, copy out
namespace {
template <typename F>
int with_host_vectors(ibim_plan *pl, size_t n_in, const double *in, size_t n_out, double *out, F fn) {
    double *d_in = nullptr, *d_out = nullptr;
    IBIM_HIP_CHECK(dev_malloc(&d_in, (n_in ? n_in : 1) * sizeof(double)));
    hipError_t e = dev_malloc(&d_out, (n_out ? n_out : 1) * sizeof(double));
    if (e != hipSuccess) {
        (void)hipFree(d_in);
        set_error("hipMalloc (host wrapper)");
        return IBIM_ERR_NOMEM;
    }
    e = hipMemcpyAsync(d_in, in, n_in * sizeof(double), hipMemcpyHostToDevice, pl->stream);
    int rc = IBIM_OK;
    if (e == hipSuccess) rc = fn(d_in, d_out);
    else rc = IBIM_ERR_HIP;
    if (rc >= 0) {
        e = hipMemcpyAsync(out, d_out, n_out * sizeof(double), hipMemcpyDeviceToHost, pl->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
        if (e != hipSuccess) {
            set_error(std::string("copy back: ") + hipGetErrorString(e));
            rc = IBIM_ERR_HIP;
        }
    } else {
        (void)hipStreamSynchronize(pl->stream);
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}
}  // namespace

namespace ibim {
namespace {
struct BlockCache {
    std::mutex mu;
    std::unordered_map<void *, std::pair<int, size_t>> live;  // block -> (device, size)
    std::map<int, std::multimap<size_t, void *>> idle;         // per device: size -> block
    std::map<int, size_t> idle_bytes;
};
BlockCache &block_cache() {
    static BlockCache *c = new BlockCache();  // kept to process exit
    return *c;
}
constexpr size_t kBlockRound = (size_t)2 << 20;
constexpr size_t kIdleCap = (size_t)64 << 30;  // idle bytes kept per device

// frees every idle cached block of device `dev`
void flush_idle(int dev) {
    BlockCache &c = block_cache();
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        for (auto &kv : c.idle[dev]) drop.push_back(kv.second);
        c.idle[dev].clear();
        c.idle_bytes[dev] = 0;
    }
    for (void *q : drop) (void)hipFree(q);
}
}  // namespace

hipError_t dev_malloc_raw(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) return e;
    (void)hipGetLastError();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return e;
    flush_idle(dev);
    return hipMalloc(p, bytes);
}

hipError_t block_alloc(void **p, size_t bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t sz = (bytes + kBlockRound - 1) / kBlockRound * kBlockRound;
    BlockCache &c = block_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto &m = c.idle[dev];
        auto it = m.lower_bound(sz);
        if (it != m.end() && it->first <= sz + sz / 4) {  // at most a quarter larger
            *p = it->second;
            c.live[*p] = {dev, it->first};
            c.idle_bytes[dev] -= it->first;
            m.erase(it);
            return hipSuccess;
        }
    }
    e = dev_malloc_raw(p, sz);  // (returns this device's idle blocks and retries on failure)
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(c.mu);
    c.live[*p] = {dev, sz};
    return hipSuccess;
}

void block_release(void *p) {
    if (!p) return;
    BlockCache &c = block_cache();
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.live.find(p);
        if (it == c.live.end()) return;  // not a block of this cache
        const int dev = it->second.first;
        const size_t sz = it->second.second;
        c.live.erase(it);
        auto &m = c.idle[dev];
        m.emplace(sz, p);
        c.idle_bytes[dev] += sz;
        while (c.idle_bytes[dev] > kIdleCap && !m.empty()) {  // drop the largest
            auto last = std::prev(m.end());
            drop.push_back(last->second);
            c.idle_bytes[dev] -= last->first;
            m.erase(last);
        }
    }
    for (void *q : drop) (void)hipFree(q);
}
}  // namespace ibim

extern "C" {

const char *ibim_last_error(void) { return g_last_error.c_str(); }
const char *ibim_version(void) { return "ibim-levelset-amd 0.1 (gfx950)"; }

void ibim_plan_destroy(ibim_plan *pl) {
    if (!pl) return;
    ibim_plan_destroy(pl->inner);  // (it runs on this plan's stream; it synchronises first)
    // both streams idle before any block returns to the shared cache (a
    // matvec that failed between its fork and join may still have near-field
    // work in flight on the second stream)
    if (pl->stream) (void)hipStreamSynchronize(pl->stream);
    if (pl->stream2) (void)hipStreamSynchronize(pl->stream2);
    for (auto &r : pl->timer.pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : pl->timer.pool) (void)hipEventDestroy(e);
    for (void *p : pl->allocs) block_release(p);
    if (pl->gm_host) (void)hipHostFree(pl->gm_host);
    pl->pool.free_all();
    pl->spool.free_all();
    if (pl->stream2) {
        (void)hipStreamSynchronize(pl->stream2);
        (void)hipStreamDestroy(pl->stream2);
    }
    if (pl->ev_fork) (void)hipEventDestroy(pl->ev_fork);
    if (pl->ev_join) (void)hipEventDestroy(pl->ev_join);
    if (pl->stream && pl->stream != pl->own_stream) (void)hipStreamSynchronize(pl->stream);
    if (pl->ev_switch) (void)hipEventDestroy(pl->ev_switch);
    if (pl->owns_stream && pl->own_stream) (void)hipStreamDestroy(pl->own_stream);
    delete pl;
}

int ibim_plan_create_device(const ibim_params *prm, int64_t n, const double *d_xyz, const double *d_nrm,
                            const double *d_w, ibim_plan **out) {
    if (!out) {
        set_error("out is NULL");
        return IBIM_ERR_ARG;
    }
    *out = nullptr;
    int rc = check_params(prm, n);
    if (rc) return rc;
    if (!d_xyz) {
        set_error("xyz is NULL");
        return IBIM_ERR_ARG;
    }
    int dev = 0;
    IBIM_HIP_CHECK(hipGetDevice(&dev));
    ibim_plan *pl = new ibim_plan();
    pl->prm = *prm;
    pl->n = (int)n;
    pl->p = prm->p;
    pl->r = prm->p * prm->p * prm->p;
    pl->device = dev;
    hipError_t e = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        delete pl;
        return IBIM_ERR_HIP;
    }
    pl->owns_stream = true;
    pl->own_stream = pl->stream;
    if ((rc = build_chebyshev(pl)) || (rc = build_tree(pl, d_xyz, d_nrm, d_w)) || (rc = build_m2l_operators(pl)) ||
        (rc = build_vpairs(pl, pl->need_local, pl->vp, nullptr)) || (rc = build_xpairs(pl))) {
        ibim_plan_destroy(pl);
        return rc;
    }
    // (with a lower inner order the mixed operators wait for this plan's first mixed matvec)
    if (pl->prm.precision != IBIM_PREC_FP64 && pl->prm.inner_order == 0 &&
        (rc = pl->prm.precision == IBIM_PREC_MIXED ? build_m2l_operators_f32(pl) : build_m2l_operators_f16(pl))) {
        ibim_plan_destroy(pl);
        return rc;
    }
    {
        TargetSet ts = source_targets(pl);
        if ((rc = build_p2p_ranges(pl, ts, pl->p2p, nullptr))) {
            ibim_plan_destroy(pl);
            return rc;
        }
        pl->p2p_roff = ts.roff;
        pl->p2p_ranges = ts.ranges;
    }
    const size_t fields = (size_t)pl->nnodes * 4 * pl->r;
    if ((rc = dalloc(pl, &pl->M, fields)) || (rc = dalloc(pl, &pl->L, fields)) ||
        (rc = dalloc(pl, &pl->Q, 4 * (size_t)pl->n)) || (rc = dalloc(pl, &pl->F, 4 * (size_t)pl->n))) {
        ibim_plan_destroy(pl);
        return rc;
    }
    e = hipStreamSynchronize(pl->stream);
    if (e != hipSuccess) {
        set_error(std::string("plan build: ") + hipGetErrorString(e));
        ibim_plan_destroy(pl);
        return IBIM_ERR_HIP;
    }
    *out = pl;
    return IBIM_OK;
}

int ibim_plan_create(const ibim_params *prm, int64_t n, const double *xyz, const double *nrm, const double *w,
                     ibim_plan **out) {
    if (!out) {
        set_error("out is NULL");
        return IBIM_ERR_ARG;
    }
    *out = nullptr;
    int rc = check_params(prm, n);
    if (rc) return rc;
    if (!xyz || !nrm || !w) {
        set_error("xyz, normals and weights are required");
        return IBIM_ERR_ARG;
    }
    double *d = nullptr;
    IBIM_HIP_CHECK(dev_malloc(&d, 7 * (size_t)n * sizeof(double)));
    hipError_t e = hipMemcpy(d, xyz, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 3 * n, nrm, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 6 * n, w, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
        return IBIM_ERR_HIP;
    }
    rc = ibim_plan_create_device(prm, n, d, d + 3 * n, d + 6 * n, out);
    (void)hipFree(d);
    return rc;
}

int ibim_plan_sync(ibim_plan *pl) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(pl->stream));
    return IBIM_OK;
}

int ibim_plan_set_stream(ibim_plan *pl, void *stream) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    hipStream_t next = stream ? (hipStream_t)stream : pl->own_stream;
    if (next == pl->stream) return IBIM_OK;
    // what the plan has queued so far comes before anything on the new stream
    if (!pl->ev_switch) IBIM_HIP_CHECK(hipEventCreateWithFlags(&pl->ev_switch, hipEventDisableTiming));
    IBIM_HIP_CHECK(hipEventRecord(pl->ev_switch, pl->stream));
    IBIM_HIP_CHECK(hipStreamWaitEvent(next, pl->ev_switch, 0));
    pl->stream = next;
    return IBIM_OK;
}

int ibim_plan_set_precision(ibim_plan *pl, int precision) {
    if (!pl || (precision != IBIM_PREC_FP64 && precision != IBIM_PREC_MIXED && precision != IBIM_PREC_MIXED_F16)) {
        set_error("bad plan or precision");
        return IBIM_ERR_ARG;
    }
    // the operators a precision needs are built before the plan switches to
    // it: a failed build leaves the plan in its previous precision.  With a
    // lower inner order the plan's own mixed operators wait for its first
    // mixed matvec (as at creation), and a companion plan already built for
    // the inner solves switches with it
    int rc = IBIM_OK;
    if (pl->prm.inner_order > 0) {
        if (pl->inner && precision != IBIM_PREC_FP64) rc = ibim_plan_set_precision(pl->inner, precision);
    } else if (precision == IBIM_PREC_MIXED) rc = build_m2l_operators_f32(pl);
    else if (precision == IBIM_PREC_MIXED_F16) rc = build_m2l_operators_f16(pl);
    if (rc == IBIM_OK) pl->prm.precision = precision;
    return rc;
}

int64_t ibim_plan_footprint(const ibim_plan *pl) {
    return pl ? pl->footprint + (pl->inner ? pl->inner->footprint : 0) : 0;
}

int ibim_plan_profile(ibim_plan *pl, int enable) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    pl->timer.on = enable != 0;
    pl->timer.mask = enable ? ~0u : 0u;
    if (pl->inner) ibim_plan_profile(pl->inner, enable);
    return IBIM_OK;
}

int ibim_plan_set_trace(ibim_plan *pl, int on) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    pl->trace = on != 0;
    if (pl->inner) pl->inner->trace = pl->trace;
    return IBIM_OK;
}

int ibim_plan_profile_stages(ibim_plan *pl, unsigned stage_mask) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    pl->timer.mask = stage_mask & ((1u << IBIM_STAGE_COUNT) - 1u);
    pl->timer.on = pl->timer.mask != 0;
    if (pl->inner) ibim_plan_profile_stages(pl->inner, stage_mask);
    return IBIM_OK;
}

int ibim_plan_stage_report(ibim_plan *pl, ibim_stage_report *out, int reset) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    int rc = timer_report(pl->timer, out, reset);
    if (rc || !pl->inner) return rc;
    // the inner plan's stages (the mixed matvecs of the inner solves) count as this plan's
    ibim_stage_report r;
    if ((rc = timer_report(pl->inner->timer, out ? &r : nullptr, reset)) || !out) return rc;
    for (int i = 0; i < IBIM_STAGE_SLOTS; ++i) {
        out->ms[i] += r.ms[i];
        out->launches[i] += r.launches[i];
        out->valu_flops[i] += r.valu_flops[i];
        out->mfma_flops[i] += r.mfma_flops[i];
        out->bytes[i] += r.bytes[i];
    }
    return IBIM_OK;
}

int ibim_plan_stage_times(ibim_plan *pl, double *ms, int64_t *launches, int reset) {
    ibim_stage_report r;
    const int rc = ibim_plan_stage_report(pl, &r, reset);
    if (rc) return rc;
    for (int i = 0; i < IBIM_STAGE_COUNT; ++i) {
        if (ms) ms[i] = r.ms[i];
        if (launches) launches[i] = r.launches[i];
    }
    return IBIM_OK;
}

int ibim_matvec_device(ibim_plan *pl, const double *d_x, double *d_y) {
    if (!pl || !d_x || !d_y) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    return matvec_device(pl, d_x, d_y, pl->prm.precision != IBIM_PREC_FP64);
}

int ibim_matvec(ibim_plan *pl, const double *x, double *y) {
    if (!pl || !x || !y) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    const size_t n2 = 2 * (size_t)pl->n;
    return with_host_vectors(pl, n2, x, n2, y, [&](const double *dx, double *dy) {
        return matvec_device(pl, dx, dy, pl->prm.precision != IBIM_PREC_FP64);
    });
}

struct ibim_group {
    ibim::MatvecGroup *g = nullptr;
};

int ibim_group_create(ibim_plan *const *plans, int k, int mixed, ibim_group **out) {
    if (!out || !plans) {
        set_error("bad argument to ibim_group_create");
        return IBIM_ERR_ARG;
    }
    *out = nullptr;
    MatvecGroup *g = nullptr;
    if (int rc = group_create(plans, k, mixed != 0, &g)) return rc;
    *out = new ibim_group{g};
    return IBIM_OK;
}

int ibim_group_matvec_device(ibim_group *g, const double *const *d_x, double *const *d_y) {
    if (!g || !d_x || !d_y) {
        set_error("bad argument to ibim_group_matvec_device");
        return IBIM_ERR_ARG;
    }
    for (int m = 0; m < g->g->k; ++m)
        if (!d_x[m] || !d_y[m]) {
            set_error("ibim_group_matvec_device: NULL vector");
            return IBIM_ERR_ARG;
        }
    IBIM_HIP_CHECK(hipSetDevice(g->g->device));
    return group_matvec_ordered(g->g, d_x, d_y);
}

int ibim_group_stage_report(ibim_group *g, ibim_stage_report *out, int reset) {
    if (!g || !out) {
        set_error("bad argument to ibim_group_stage_report");
        return IBIM_ERR_ARG;
    }
    return timer_report(g->g->timer, out, reset);
}

void ibim_group_destroy(ibim_group *g) {
    if (!g) return;
    group_destroy(g->g);
    delete g;
}

int ibim_kernel_fmm(ibim_plan *pl, int kernel_id, const double *charge, double *potential) {
    if (!pl || !charge || !potential) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    const size_t n = (size_t)pl->n;
    return with_host_vectors(pl, n, charge, n, potential,
                             [&](const double *dq, double *dp) { return kernel_fmm_device(pl, kernel_id, dq, dp); });
}

int ibim_kernel_fmm_targets(ibim_plan *pl, int kernel_id, int64_t m, const double *txyz, const double *charge,
                            double *potential) {
    if (!pl || m < 1 || !txyz || !charge || !potential || kernel_id < 0 || kernel_id > 15) {
        set_error("bad argument to ibim_kernel_fmm_targets");
        return IBIM_ERR_ARG;
    }
    const size_t n = (size_t)pl->n;
    std::vector<double> packed(3 * (size_t)m + n);
    std::memcpy(packed.data(), txyz, 3 * (size_t)m * sizeof(double));
    std::memcpy(packed.data() + 3 * m, charge, n * sizeof(double));
    return with_host_vectors(pl, packed.size(), packed.data(), (size_t)m, potential, [&](const double *d, double *dp) {
        return target_fmm_device(pl, kernel_id, (int)m, d, d + 3 * m, dp);
    });
}

int ibim_rhs_device(ibim_plan *pl, int64_t m, const double *d_axyz, const double *d_aq, double *d_b) {
    if (!pl || m < 0 || (m > 0 && (!d_axyz || !d_aq)) || !d_b) {
        set_error("bad argument to ibim_rhs");
        return IBIM_ERR_ARG;
    }
    return rhs_device(pl, (int)m, d_axyz, d_aq, d_b);
}

int ibim_rhs(ibim_plan *pl, int64_t m, const double *axyz, const double *aq, double *b) {
    if (!pl || m < 0 || (m > 0 && (!axyz || !aq)) || !b) {
        set_error("bad argument to ibim_rhs");
        return IBIM_ERR_ARG;
    }
    std::vector<double> packed(4 * (size_t)m);
    if (m > 0) {
        std::memcpy(packed.data(), axyz, 3 * (size_t)m * sizeof(double));
        std::memcpy(packed.data() + 3 * m, aq, (size_t)m * sizeof(double));
    }
    return with_host_vectors(pl, packed.size(), packed.data(), 2 * (size_t)pl->n, b,
                             [&](const double *d, double *db) { return rhs_device(pl, (int)m, d, d + 3 * m, db); });
}

int ibim_polarize_device(ibim_plan *pl, int64_t m, const double *d_axyz, const double *d_x, double *d_phi) {
    if (!pl || m < 1 || !d_axyz || !d_x || !d_phi) {
        set_error("bad argument to ibim_polarize");
        return IBIM_ERR_ARG;
    }
    return polarize_device(pl, (int)m, d_axyz, d_x, d_phi);
}

int ibim_polarize_device_noalloc(ibim_plan *pl, int64_t m, const double *d_axyz, const double *d_x, double *d_phi) {
    if (!pl || m < 1 || m > (int64_t)0x7fffffff || !d_axyz || !d_x || !d_phi) {
        set_error("bad argument to ibim_polarize_device_noalloc");
        return IBIM_ERR_ARG;
    }
    return polarize_noalloc(pl, (int)m, d_axyz, d_x, d_phi);
}

int ibim_polarize(ibim_plan *pl, int64_t m, const double *axyz, const double *x, double *phi) {
    if (!pl || m < 1 || !axyz || !x || !phi) {
        set_error("bad argument to ibim_polarize");
        return IBIM_ERR_ARG;
    }
    const size_t n2 = 2 * (size_t)pl->n;
    std::vector<double> packed(3 * (size_t)m + n2);
    std::memcpy(packed.data(), axyz, 3 * (size_t)m * sizeof(double));
    std::memcpy(packed.data() + 3 * m, x, n2 * sizeof(double));
    return with_host_vectors(pl, packed.size(), packed.data(), (size_t)m, phi, [&](const double *d, double *dphi) {
        return polarize_device(pl, (int)m, d, d + 3 * m, dphi);
    });
}

int ibim_gmres_device(ibim_plan *pl, const double *d_b, double *d_x, int restart, int maxit, double tol,
                      ibim_solve_info *info) {
    if (!pl || !d_b || !d_x) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    return gmres_device(pl, d_b, d_x, restart, maxit, tol, info);
}

int ibim_gmres(ibim_plan *pl, const double *b, double *x, int restart, int maxit, double tol, ibim_solve_info *info) {
    if (!pl || !b || !x) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    const size_t n2 = 2 * (size_t)pl->n;
    std::vector<double> packed(2 * n2);
    std::memcpy(packed.data(), b, n2 * sizeof(double));
    std::memcpy(packed.data() + n2, x, n2 * sizeof(double));
    return with_host_vectors(pl, packed.size(), packed.data(), n2, x, [&](const double *d, double *dx) {
        hipError_t e = hipMemcpyAsync(dx, d + n2, n2 * sizeof(double), hipMemcpyDeviceToDevice, pl->stream);
        if (e != hipSuccess) {
            set_error("hipMemcpy (gmres x0)");
            return IBIM_ERR_HIP;
        }
        return gmres_device(pl, d, dx, restart, maxit, tol, info);
    });
}

// ------------------------------------------------------------ inspection --
int ibim_tree_export(ibim_plan *pl, int *level, int *parent, int *slot, int *leaf, int *empty, int *nsrc,
                     double *centre, double *radius, int *nu, int *nv, int *nw, int *nx) {
    if (!pl) {
        set_error("plan is NULL");
        return IBIM_ERR_ARG;
    }
    const int nn = pl->nnodes;
    std::vector<int> lv(nn), pa(nn), sl(nn), fl(nn), sb(nn), se(nn);
    std::vector<int> off[4];
    std::vector<double> c(3 * (size_t)nn);
    IBIM_HIP_CHECK(hipStreamSynchronize(pl->stream));
    IBIM_HIP_CHECK(hipMemcpy(lv.data(), pl->level, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(pa.data(), pl->parent, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(sl.data(), pl->slot, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(fl.data(), pl->flags, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(sb.data(), pl->sbeg, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(se.data(), pl->send, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(c.data(), pl->ctr, 3 * (size_t)nn * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; ++k) {
        off[k].resize(nn + 1);
        IBIM_HIP_CHECK(hipMemcpy(off[k].data(), pl->loff[k], (nn + 1) * sizeof(int), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nn; ++i) {
        if (level) level[i] = lv[i];
        if (parent) parent[i] = pa[i];
        if (slot) slot[i] = sl[i];
        if (leaf) leaf[i] = (fl[i] & kLeaf) ? 1 : 0;
        if (empty) empty[i] = (fl[i] & kEmpty) ? 1 : 0;
        if (nsrc) nsrc[i] = se[i] - sb[i];
        for (int d = 0; d < 3; ++d) {
            if (centre) centre[3 * i + d] = c[3 * (size_t)i + d];
            if (radius) radius[3 * i + d] = pl->lvl_rad[3 * lv[i] + d];
        }
        if (nu) nu[i] = off[0][i + 1] - off[0][i];
        if (nv) nv[i] = off[1][i + 1] - off[1][i];
        if (nw) nw[i] = off[2][i + 1] - off[2][i];
        if (nx) nx[i] = off[3][i + 1] - off[3][i];
    }
    return IBIM_OK;
}

int ibim_tree_lists(ibim_plan *pl, int kind, int *offsets, int *items, int64_t cap) {
    if (!pl || kind < 0 || kind > 3 || !offsets) {
        set_error("bad argument to ibim_tree_lists");
        return IBIM_ERR_ARG;
    }
    IBIM_HIP_CHECK(hipStreamSynchronize(pl->stream));
    IBIM_HIP_CHECK(hipMemcpy(offsets, pl->loff[kind], (pl->nnodes + 1) * sizeof(int), hipMemcpyDeviceToHost));
    const int64_t tot = pl->ltot[kind];
    if (items) {
        if (cap < tot) {
            set_error("items capacity too small");
            return IBIM_ERR_ARG;
        }
        if (tot > 0) IBIM_HIP_CHECK(hipMemcpy(items, pl->litem[kind], tot * sizeof(int), hipMemcpyDeviceToHost));
    }
    return IBIM_OK;
}

int ibim_tree_points(ibim_plan *pl, int id, int *out, int64_t cap) {
    if (!pl || id < 0 || id >= pl->nnodes) {
        set_error("bad node id");
        return IBIM_ERR_ARG;
    }
    int b = 0, e = 0;
    IBIM_HIP_CHECK(hipStreamSynchronize(pl->stream));
    IBIM_HIP_CHECK(hipMemcpy(&b, pl->sbeg + id, sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(&e, pl->send + id, sizeof(int), hipMemcpyDeviceToHost));
    const int cnt = e - b;
    if (out && cnt > 0) {
        if (cap < cnt) {
            set_error("capacity too small");
            return IBIM_ERR_ARG;
        }
        IBIM_HIP_CHECK(hipMemcpy(out, pl->perm + b, cnt * sizeof(int), hipMemcpyDeviceToHost));
    }
    return cnt;
}

int ibim_tree_stats_get(ibim_plan *pl, ibim_tree_stats *st) {
    if (!pl || !st) {
        set_error("null argument");
        return IBIM_ERR_ARG;
    }
    const int nn = pl->nnodes;
    std::vector<int> fl(nn), sb(nn), se(nn);
    std::vector<int> off[4], it[4];
    IBIM_HIP_CHECK(hipStreamSynchronize(pl->stream));
    IBIM_HIP_CHECK(hipMemcpy(fl.data(), pl->flags, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(sb.data(), pl->sbeg, nn * sizeof(int), hipMemcpyDeviceToHost));
    IBIM_HIP_CHECK(hipMemcpy(se.data(), pl->send, nn * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; ++k) {
        off[k].resize(nn + 1);
        it[k].resize(pl->ltot[k] > 0 ? pl->ltot[k] : 1);
        IBIM_HIP_CHECK(hipMemcpy(off[k].data(), pl->loff[k], (nn + 1) * sizeof(int), hipMemcpyDeviceToHost));
        if (pl->ltot[k] > 0)
            IBIM_HIP_CHECK(hipMemcpy(it[k].data(), pl->litem[k], pl->ltot[k] * sizeof(int), hipMemcpyDeviceToHost));
    }
    ibim_tree_stats s{};
    s.n_points = pl->n;
    s.n_nodes = nn;
    s.depth = pl->depth;
    for (int k = 0; k < 4; ++k) (&s.n_u)[k] = pl->ltot[k];
    for (int i = 0; i < nn; ++i) {
        const bool leaf = fl[i] & kLeaf, empty = fl[i] & kEmpty;
        if (leaf) s.n_leaves++;
        if (leaf && empty) s.n_empty_leaves++;
        const int64_t nt = se[i] - sb[i];
        if (leaf && nt > 0) {
            int64_t ns = 0;
            for (int k : {0, 2}) {
                int64_t nk = 0;
                for (int q = off[k][i]; q < off[k][i + 1]; ++q) nk += se[it[k][q]] - sb[it[k][q]];
                ns += nk;
                if (k == 0) s.p2p_pairs_u += nt * nk;
            }
            s.p2p_pairs += nt * ns;
        }
        if (i != 0 && nt > 0)
            for (int k : {1, 3})
                for (int q = off[k][i]; q < off[k][i + 1]; ++q)
                    if (!(fl[it[k][q]] & kEmpty)) {
                        s.m2l_pairs++;
                        if (k == 1) s.v_pairs++;
                        else s.x_pairs++;
                    }
    }
    *st = s;
    return IBIM_OK;
}

}  // extern "C"
