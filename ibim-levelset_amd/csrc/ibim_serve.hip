// ibim_serve.hip -- continuous batching of PB solves on one GPU.
//
// Workers (one host thread per slot) take requests from a queue kept in
// submission order.  A free worker starts the first queued request whose
// estimated device footprint fits beside the running requests within the
// memory budget, so a small request may pass a large one that does not fit
// yet; a request passed over kMaxPassed times holds back everything behind
// it until it runs (no starvation), and a request that could never fit runs
// alone.  Each request builds its own level set, tube nodes and plan on the
// worker's stream, so requests of different sizes overlap on the device;
// everything a solve needs stays in HBM until it finishes.
//
// Fusion: the GMRES matvecs of small requests (at most fuse_max tube
// points) go through a batcher.  A solve that needs a matvec waits until
// every registered solve waits too (they fall into lockstep: each GMRES
// iteration is one matvec plus short orthogonalisation steps), or until
// fuse_wait has passed, and then all waiting matvecs of one form run as one
// fused group matvec: one launch per FMM stage over all of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "ibim_internal.h"

namespace ibim {

struct Batcher {
    using clock = std::chrono::steady_clock;
    struct Req {
        ibim_plan *pl;
        const double *x;
        double *y;
        bool mixed;
        int rc = IBIM_OK;
        bool done = false;
    };
    struct Cached {
        std::vector<ibim_plan *> plans;  // in the group's member order
        bool mixed;
        MatvecGroup *g;
        uint64_t used;
    };
    std::mutex mu;  // every field; the batch that completes the set is launched under it
    std::condition_variable cv;
    int active = 0;                      // registered solves
    std::vector<Req *> pending;
    std::vector<Cached> cache;           // groups by member set
    uint64_t tick = 0;
    std::chrono::microseconds wait{2000};
    int64_t group_matvecs = 0, member_matvecs = 0;
    ibim_stage_report acc{};             // stage reports of released groups
    static constexpr size_t kMaxCached = 8;

    void add_report(MatvecGroup *g) {
        ibim_stage_report r;
        if (timer_report(g->timer, &r, 1) != IBIM_OK) return;
        for (int i = 0; i < IBIM_STAGE_SLOTS; ++i) {
            acc.ms[i] += r.ms[i];
            acc.launches[i] += r.launches[i];
            acc.valu_flops[i] += r.valu_flops[i];
            acc.mfma_flops[i] += r.mfma_flops[i];
            acc.bytes[i] += r.bytes[i];
        }
    }
    void release(size_t i) {
        add_report(cache[i].g);
        group_destroy(cache[i].g);
        cache.erase(cache.begin() + (long)i);
    }
    // the group of these plans (sorted), built on first use; nullptr if they
    // cannot run as one group (the caller runs them one by one)
    MatvecGroup *group_for(const std::vector<ibim_plan *> &plans, bool mixed) {
        for (Cached &c : cache)
            if (c.mixed == mixed && c.plans == plans) {
                c.used = ++tick;
                return c.g;
            }
        MatvecGroup *g = nullptr;
        if (group_create(plans.data(), (int)plans.size(), mixed, &g) != IBIM_OK) return nullptr;
        if (cache.size() >= kMaxCached) {
            size_t lru = 0;
            for (size_t i = 1; i < cache.size(); ++i)
                if (cache[i].used < cache[lru].used) lru = i;
            release(lru);
        }
        cache.push_back({plans, mixed, g, ++tick});
        return g;
    }
    // run every pending matvec (under mu): per form, one fused group of all
    // its requests, or single matvecs where a group cannot be formed
    void launch_pending() {
        std::vector<Req *> batch;
        batch.swap(pending);
        for (int form = 0; form < 2; ++form) {
            std::vector<Req *> part;
            for (Req *r : batch)
                if (r->mixed == (form == 1)) part.push_back(r);
            for (size_t i0 = 0; i0 < part.size(); i0 += kMaxGroup) {
                const size_t i1 = std::min(part.size(), i0 + (size_t)kMaxGroup);
                std::vector<Req *> chunk(part.begin() + (long)i0, part.begin() + (long)i1);
                std::sort(chunk.begin(), chunk.end(), [](Req *a, Req *b) { return a->pl < b->pl; });
                MatvecGroup *g = nullptr;
                if (chunk.size() > 1) {
                    std::vector<ibim_plan *> plans;
                    for (Req *r : chunk) plans.push_back(r->pl);
                    g = group_for(plans, form == 1);
                }
                if (g) {
                    std::vector<const double *> xs;
                    std::vector<double *> ys;
                    for (Req *r : chunk) {
                        xs.push_back(r->x);
                        ys.push_back(r->y);
                    }
                    const int rc = group_matvec_ordered(g, xs.data(), ys.data());
                    for (Req *r : chunk) r->rc = rc;
                    group_matvecs++;
                    member_matvecs += (int64_t)chunk.size();
                } else {
                    for (Req *r : chunk) r->rc = matvec_single(r->pl, r->x, r->y, r->mixed);
                }
            }
        }
        for (Req *r : batch) r->done = true;
    }
    void enter(ibim_plan *pl) {
        std::lock_guard<std::mutex> lk(mu);
        pl->batcher = this;
        active++;
    }
    void leave(ibim_plan *pl) {
        std::lock_guard<std::mutex> lk(mu);
        pl->batcher = nullptr;
        active--;
        for (size_t i = cache.size(); i-- > 0;)
            if (std::find(cache[i].plans.begin(), cache[i].plans.end(), pl) != cache[i].plans.end()) release(i);
        cv.notify_all();  // the waiting solves may now be all there is
    }
    ~Batcher() {
        while (!cache.empty()) release(cache.size() - 1);
    }
};

int batch_matvec(Batcher *b, ibim_plan *pl, const double *d_x, double *d_y, bool mixed) {
    Batcher::Req r{pl, d_x, d_y, mixed};
    std::unique_lock<std::mutex> lk(b->mu);
    b->pending.push_back(&r);
    const auto deadline = Batcher::clock::now() + b->wait;
    while (!r.done) {
        if ((int)b->pending.size() >= b->active || Batcher::clock::now() >= deadline) {
            b->launch_pending();
            b->cv.notify_all();
        } else {
            b->cv.wait_until(lk, deadline);
        }
    }
    return r.rc;
}

}  // namespace ibim

struct ibim_server {
    struct Job {
        int64_t id;
        ibim_request req;
        std::vector<double> xyz, rad, q;
        int64_t est = 0;
        int passed = 0;  // times a later request was started before this one
        ibim_result res{};
        bool done = false;
    };
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::map<int64_t, Job> jobs;
    std::vector<std::thread> workers;
    int64_t next_id = 0;
    int64_t budget = 0, in_use = 0;
    int running = 0;
    bool stop = false;
    int device = 0;
    unsigned stage_mask = 0;    // stages recorded in each request's plan
    ibim_stage_report acc{};    // summed over finished requests
    int64_t fuse_max = 0;       // requests of at most this many tube points fuse their matvecs
    ibim::Batcher batcher;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace ibim {
namespace {

// device bytes a request will hold: the level-set grids (4 arrays while the
// surface is built) or the plan, whichever is larger, plus the tube nodes.
// Tube points are bounded by the atoms' probe-expanded sphere area times the
// tube's (2 thickness) layers per unit area, capped by the grid's surface;
// each point costs ~ (4 r fields x 2 expansions per ~20 points per box +
// vectors + GMRES basis) bytes.
int64_t estimate_footprint(const ibim_request &r, const std::vector<double> &xyz, const std::vector<double> &rad) {
    const double h = r.geom.h;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, area = 0.0;
    for (int64_t a = 0; a < r.n_atoms; ++a) {
        const double rr = rad[a] + r.geom.probe;
        area += 4.0 * M_PI * rr * rr;
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], xyz[3 * a + d] - rr);
            hi[d] = std::max(hi[d], xyz[3 * a + d] + rr);
        }
    }
    double ext[3], grid = 1.0;
    for (int d = 0; d < 3; ++d) {
        ext[d] = hi[d] - lo[d] + 2 * r.geom.pad;
        grid *= ext[d] / h + 1;
    }
    const double box_area = 2 * (ext[0] * ext[1] + ext[1] * ext[2] + ext[0] * ext[2]);
    const double npts = std::min(area, 3.0 * box_area) * (2.0 * r.geom.thickness) / (h * h);
    // one plan of order p: expansions, vectors, the GMRES basis if it holds
    // it, its operators and (FP16 form) its pair rows
    auto plan_bytes = [&](int p, bool basis) {
        const int rr3 = p * p * p;
        const double per_point = 2.0 * 4 * rr3 * 8 / 12.0 + 8.0 * ((basis ? 2 * (r.restart + 1) : 6) + 40);
        const double ops = (r.fmm.precision != IBIM_PREC_FP64 ? 1.5 : 1.0) * 12 * 56 * pow(4.0 * rr3, 2) * 8;
        // the FP16 V list keeps one FP24 row of 4 p^3 (rounded up to 32) per V
        // interaction; about 340 / leaf capacity V interactions per tube point
        // (plus the X list's, a few per cent), and one unfolded FP16 operator
        // per (class, mask) key with pairs, V and X (a few thousand keys)
        const double yv = (r.fmm.precision == IBIM_PREC_MIXED_F16 && p <= 6)
                              ? npts * 350.0 / std::max(1, r.fmm.max_points) * ((4 * rr3 + 31) / 32 * 32) * 3 +
                                    2500.0 * pow(4.0 * rr3, 2) * 2
                              : 0.0;
        return npts * per_point + ops + yv;
    };
    // a mixed solve with a lower inner order also holds the inner plan, which
    // then keeps the Krylov basis
    const bool two = r.fmm.precision != IBIM_PREC_FP64 && r.fmm.inner_order > 0;
    const double plans = plan_bytes(r.fmm.p, !two) + (two ? plan_bytes(r.fmm.inner_order, true) : 0.0);
    return (int64_t)(std::max(grid * 8 * 4.2, plans) + npts * 56);
}

// a request passed over this many times is started before any later one
constexpr int kMaxPassed = 8;

void run_job(ibim_server *s, ibim_server::Job &j, hipStream_t st) {
    ibim_result &res = j.res;
    ibim_surface *surf = nullptr;
    ibim_plan *plan = nullptr;
    double *d_b = nullptr, *d_x = nullptr, *d_at = nullptr, *d_phi = nullptr;
    const int64_t m = j.req.n_atoms;
    int rc = ibim_surface_from_atoms(&j.req.geom, m, j.xyz.data(), j.rad.data(), &surf);
    const double *sx = nullptr, *sn = nullptr, *sw = nullptr;
    ibim_surface_info inf{};
    if (rc == IBIM_OK) rc = ibim_surface_get_info(surf, &inf);
    if (rc == IBIM_OK) rc = ibim_surface_device(surf, &sx, &sn, &sw);
    int64_t fuse_max = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        fuse_max = s->fuse_max;
    }
    // a solve whose matvecs go through the batcher applies this plan's own
    // operator in its inner solves (a second plan's would bypass the batch)
    const bool fuse = rc == IBIM_OK && fuse_max > 0 && inf.n_nodes <= fuse_max;
    ibim_params fmm = j.req.fmm;
    if (fuse) fmm.inner_order = 0;
    if (rc == IBIM_OK) rc = ibim_plan_create_device(&fmm, inf.n_nodes, sx, sn, sw, &plan);
    const size_t n2 = 2 * (size_t)inf.n_nodes;
    unsigned mask = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        mask = s->stage_mask;
    }
    if (rc == IBIM_OK && mask) rc = ibim_plan_profile_stages(plan, mask);
    if (rc == IBIM_OK) {
        if (dev_malloc(&d_b, n2 * sizeof(double)) != hipSuccess || dev_malloc(&d_x, n2 * sizeof(double)) != hipSuccess ||
            dev_malloc(&d_at, 4 * (size_t)m * sizeof(double)) != hipSuccess ||
            dev_malloc(&d_phi, (size_t)m * sizeof(double)) != hipSuccess) {
            set_error("hipMalloc (request vectors)");
            rc = IBIM_ERR_NOMEM;
        }
    }
    if (rc == IBIM_OK) {
        std::vector<double> packed(4 * (size_t)m);
        std::copy(j.xyz.begin(), j.xyz.end(), packed.begin());
        std::copy(j.q.begin(), j.q.end(), packed.begin() + 3 * m);
        if (hipMemcpyAsync(d_at, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice, st) !=
                hipSuccess ||
            hipMemsetAsync(d_x, 0, n2 * sizeof(double), st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            rc = IBIM_ERR_HIP;
    }
    if (rc == IBIM_OK) rc = ibim_rhs_device(plan, m, d_at, d_at + 3 * m, d_b);
    const bool entered = fuse && rc == IBIM_OK;
    if (entered) s->batcher.enter(plan);
    if (rc == IBIM_OK) rc = ibim_gmres_device(plan, d_b, d_x, j.req.restart, j.req.maxit, j.req.tol, &res.info);
    if (entered) s->batcher.leave(plan);
    const int solve_rc = rc;
    if (rc >= 0) rc = ibim_polarize_device(plan, m, d_at, d_x, d_phi);
    if (rc >= 0) {
        std::vector<double> phi((size_t)m);
        if (hipMemcpyAsync(phi.data(), d_phi, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = IBIM_ERR_HIP;
        } else {
            double e = 0.0;
            for (int64_t a = 0; a < m; ++a) e += phi[a] * j.q[a];
            res.energy = 0.5 * e;
            rc = solve_rc;
        }
    }
    res.status = rc;
    res.n_points = inf.n_nodes;
    res.footprint = (plan ? ibim_plan_footprint(plan) : 0) + (int64_t)(7 * (size_t)inf.n_nodes * sizeof(double));
    ibim_stage_report rep;
    if (plan && mask && ibim_plan_stage_report(plan, &rep, 1) == IBIM_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int i = 0; i < IBIM_STAGE_SLOTS; ++i) {
            s->acc.ms[i] += rep.ms[i];
            s->acc.launches[i] += rep.launches[i];
            s->acc.valu_flops[i] += rep.valu_flops[i];
            s->acc.mfma_flops[i] += rep.mfma_flops[i];
            s->acc.bytes[i] += rep.bytes[i];
        }
    }
    (void)hipStreamSynchronize(st);
    for (double *p : {d_b, d_x, d_at, d_phi})
        if (p) (void)hipFree(p);
    ibim_plan_destroy(plan);
    ibim_surface_destroy(surf);
}

// the queued request to start next: the first one that fits, unless a
// request before it has been passed over kMaxPassed times (then none until
// that one fits); with nothing running, the first request always fits
std::deque<int64_t>::iterator next_request(ibim_server *s) {
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        const ibim_server::Job &j = s->jobs[*it];
        if (s->running == 0 || s->in_use + j.est <= s->budget) return it;
        if (j.passed >= kMaxPassed) break;
    }
    return s->queue.end();
}

void worker(ibim_server *s) {
    (void)hipSetDevice(s->device);
    hipStream_t st = nullptr;  // the worker's stream for its requests' copies
    const bool have_stream = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
    for (;;) {
        int64_t id = -1;
        ibim_server::Job *job = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                auto it = s->queue.end();
                if (!s->queue.empty()) it = next_request(s);
                if (it != s->queue.end()) {
                    for (auto q = s->queue.begin(); q != it; ++q) s->jobs[*q].passed++;
                    id = *it;
                    s->queue.erase(it);
                    job = &s->jobs[id];
                    s->in_use += job->est;
                    s->running++;
                    break;
                }
                if (s->stop && s->queue.empty()) {
                    if (have_stream) (void)hipStreamDestroy(st);
                    return;
                }
                s->cv_work.wait(lk);
            }
            job->res.t_start = s->now();
        }
        if (have_stream) {
            run_job(s, *job, st);
        } else {
            job->res.status = IBIM_ERR_HIP;
            set_error("worker stream could not be created");
        }
        {
            std::lock_guard<std::mutex> lk(s->mu);
            job->res.t_end = s->now();
            job->done = true;
            s->in_use -= job->est;
            s->running--;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
        (void)id;
    }
}

}  // namespace
}  // namespace ibim

using namespace ibim;

extern "C" {

int ibim_server_create(int slots, int64_t mem_budget, ibim_server **out) {
    if (!out || slots < 1 || slots > 64 || mem_budget < 0) {
        set_error("bad argument to ibim_server_create (1..64 slots)");
        return IBIM_ERR_ARG;
    }
    ibim_server *s = new ibim_server();
    IBIM_HIP_CHECK(hipGetDevice(&s->device));
    size_t fr = 0, tot = 0;
    IBIM_HIP_CHECK(hipMemGetInfo(&fr, &tot));
    s->budget = mem_budget > 0 ? mem_budget : (int64_t)(0.9 * (double)fr);
    for (int i = 0; i < slots; ++i) s->workers.emplace_back(worker, s);
    *out = s;
    return IBIM_OK;
}

int ibim_server_submit(ibim_server *s, const ibim_request *r, int64_t *id) {
    if (!s || !r || !id || r->n_atoms < 1 || !r->atom_xyz || !r->radii || !r->charges) {
        set_error("bad argument to ibim_server_submit");
        return IBIM_ERR_ARG;
    }
    ibim_server::Job j;
    j.req = *r;
    j.xyz.assign(r->atom_xyz, r->atom_xyz + 3 * r->n_atoms);
    j.rad.assign(r->radii, r->radii + r->n_atoms);
    j.q.assign(r->charges, r->charges + r->n_atoms);
    j.req.atom_xyz = nullptr;
    j.req.radii = nullptr;
    j.req.charges = nullptr;
    j.est = estimate_footprint(j.req, j.xyz, j.rad);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j.id = s->next_id++;
        j.res.t_submit = s->now();
        j.res.footprint_est = j.est;
        *id = j.id;
        s->queue.push_back(j.id);
        s->jobs.emplace(j.id, std::move(j));
    }
    s->cv_work.notify_all();
    return IBIM_OK;
}

int ibim_server_wait(ibim_server *s, int64_t id, ibim_result *out) {
    if (!s || !out) {
        set_error("bad argument to ibim_server_wait");
        return IBIM_ERR_ARG;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
        set_error("unknown request id");
        return IBIM_ERR_ARG;
    }
    s->cv_done.wait(lk, [&] { return it->second.done; });
    *out = it->second.res;
    s->jobs.erase(it);
    return IBIM_OK;
}

double ibim_server_now(ibim_server *s) { return s ? s->now() : 0.0; }

int ibim_server_profile_stages(ibim_server *s, unsigned stage_mask) {
    if (!s) {
        set_error("server is NULL");
        return IBIM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->stage_mask = stage_mask & ((1u << IBIM_STAGE_COUNT) - 1u);
    return IBIM_OK;
}

int ibim_server_stage_report(ibim_server *s, ibim_stage_report *out, int reset) {
    if (!s || !out) {
        set_error("bad argument to ibim_server_stage_report");
        return IBIM_ERR_ARG;
    }
    // the fused launches' stage intervals: released groups, then the live ones
    ibim_stage_report fused;
    {
        ibim::Batcher &b = s->batcher;
        std::lock_guard<std::mutex> lk(b.mu);
        for (auto &c : b.cache) b.add_report(c.g);
        fused = b.acc;
        if (reset) b.acc = ibim_stage_report{};
    }
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->acc;
    for (int i = 0; i < IBIM_STAGE_SLOTS; ++i) {
        out->ms[i] += fused.ms[i];
        out->launches[i] += fused.launches[i];
        out->valu_flops[i] += fused.valu_flops[i];
        out->mfma_flops[i] += fused.mfma_flops[i];
        out->bytes[i] += fused.bytes[i];
    }
    if (reset) s->acc = ibim_stage_report{};
    return IBIM_OK;
}

int ibim_server_set_fusion(ibim_server *s, int64_t max_points, int wait_us) {
    if (!s || max_points < 0 || wait_us < 0 || wait_us > 1000000) {
        set_error("bad argument to ibim_server_set_fusion (wait 0..1e6 us)");
        return IBIM_ERR_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->fuse_max = max_points;
    }
    std::lock_guard<std::mutex> lk(s->batcher.mu);
    s->batcher.wait = std::chrono::microseconds(wait_us);
    return IBIM_OK;
}

int ibim_server_fusion_stats(ibim_server *s, int64_t *group_matvecs, int64_t *member_matvecs) {
    if (!s) {
        set_error("server is NULL");
        return IBIM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(s->batcher.mu);
    if (group_matvecs) *group_matvecs = s->batcher.group_matvecs;
    if (member_matvecs) *member_matvecs = s->batcher.member_matvecs;
    return IBIM_OK;
}

void ibim_server_destroy(ibim_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    delete s;
}

}  // extern "C"
