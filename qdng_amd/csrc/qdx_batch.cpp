// Host runtime of the batched-jobs path (include/qdx.h, "batched jobs"):
// device arena, job admission by footprint, the job/tile tables the batched
// kernels read, and the per-tick launch sequence.
#include "qdx_batch.hpp"
#include "qdx_host_internal.hpp"

#include <deque>

namespace {

// First-fit allocator over one device allocation; blocks are 256-byte aligned
// and neighbouring free blocks are merged on release.
struct Arena {
    DevBuf buf;
    size_t cap = 0, used = 0;
    std::map<size_t, size_t> free_;   // offset -> size
    static size_t round(size_t n) { return (n + 255) & ~size_t(255); }
    void init(size_t bytes) {
        bytes = round(bytes);
        buf.alloc(bytes);
        cap = bytes;
        used = 0;
        free_.clear();
        free_[0] = bytes;
    }
    bool alloc(size_t n, size_t &off) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t sz = it->second;
            free_.erase(it);
            if (sz > n) free_[off + n] = sz - n;
            used += n;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = round(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *base() const { return buf.as<char>(); }
};

// Sub-buffers of one job inside its arena block.
struct JobLayout {
    size_t psi, A, B, X, V, mu, tk0, tk1, laser, coef, part, ctl, stats, bytes;
};

// Tiling of one job: row tiles of rb lines (NT threads), column tiles of W
// columns (256 threads).
struct JobTiling {
    int NT = 256, rb = 1, nrt = 1, W = 1, nct = 0;
};

struct BatchJob {
    int64_t id = 0;
    Grid grid;
    ChebyParams cp;
    int nsteps = 0;
    int64_t start_tick = 0;
    size_t off = 0;
    JobLayout lay{};
    JobTiling tile;
    bool has_V = false, has_mu = false, has_T = false;
    double lscale = 1.0;
    int nt = 0;
    bool active_at(int64_t t) const { return t >= start_tick && t < start_tick + nsteps; }
};

JobTiling job_tiling(const Grid &g) {
    JobTiling t;
    const int L = g.N[0], TL = L / 8;
    t.NT = TL <= 256 ? 256 : 512;
    if (g.nd == 1) {
        t.rb = 1;
        t.nrt = 1;
        return t;
    }
    t.rb = std::max(1, std::min(g.N[1], t.NT / TL));
    t.nrt = (g.N[1] + t.rb - 1) / t.rb;
    const int TL1 = g.N[1] / 8;
    int W = 1;
    while (W * 2 <= 16 && W * 2 * TL1 <= 256 && g.N[0] % (W * 2) == 0) W *= 2;
    t.W = W;
    t.nct = g.N[0] / W;
    return t;
}

JobLayout job_layout(const Grid &g, const JobTiling &t, int order, int nt, size_t rsz, bool has_mu) {
    JobLayout l{};
    size_t o = 0;
    auto take = [&](size_t n) {
        const size_t at = o;
        o += Arena::round(std::max<size_t>(n, 1));
        return at;
    };
    const size_t n = (size_t)g.size();
    l.psi = take(16 * n);
    l.A = take(2 * rsz * n);
    l.B = take(2 * rsz * n);
    l.X = g.nd == 2 ? take(2 * rsz * n) : 0;
    l.V = take(rsz * n);
    l.mu = has_mu ? take(rsz * n) : 0;
    l.tk0 = take(8 * (size_t)g.N[0]);
    l.tk1 = g.nd == 2 ? take(8 * (size_t)g.N[1]) : 0;
    l.laser = take(8 * (size_t)std::max(nt, 1));
    l.coef = take(16 * (size_t)order);
    l.part = take(8 * 3 * (size_t)t.nrt);
    l.ctl = take(sizeof(ChebyCtl));
    l.stats = take(2 * sizeof(long long));
    l.bytes = o;
    return l;
}

size_t row_lds(const Grid &g, const JobTiling &t, size_t esz, bool tw) {
    const int L = g.N[0];
    return ((size_t)2 * t.rb * (L + L / 8) + (tw ? L : 0)) * esz;
}
size_t col_lds(const Grid &g, const JobTiling &t, size_t esz, bool tw) {
    const int L = g.N[1];
    return ((size_t)2 * t.W * L + (tw ? L : 0)) * esz;
}

constexpr size_t kLds = 160 * 1024;

// Host arrays of one job from its Hamiltonian (OSum of grid terms).
struct JobHost {
    std::vector<double> V, mu, tk0, tk1, laser;
    double lscale = 1.0;
    bool has_T = false;
};

JobHost job_host(qdx_ham *h, int nsteps) {
    JobHost j;
    const Grid &g = h->grid;
    const long n = g.size();
    const std::vector<double> *las = nullptr;
    for (const Term &t : h->terms) {
        if (t.row != 0 || t.col != 0) fail(QDX_ERR_INCOMPAT, "batched jobs hold one state");
        if (t.kind == T_POT) {
            if (j.V.empty()) j.V.assign(n, 0.0);
            for (long x = 0; x < n; ++x) j.V[x] += t.v[x];
        } else if (t.kind == T_DIP) {
            if (las && (*las != t.laser || j.lscale != t.scale))
                fail(QDX_ERR_INCOMPAT, "two dipole terms with different fields on one element");
            if (j.mu.empty()) j.mu.assign(n, 0.0);
            for (long x = 0; x < n; ++x) j.mu[x] += t.v[x];
            las = &t.laser;
            j.lscale = t.scale;
        } else {
            for (int d = 0; d < g.nd; ++d) {
                std::vector<double> &dst = d == 0 ? j.tk0 : j.tk1;
                const std::vector<double> k1 = kspace_d2(g, d, t.mass[d]);
                if (dst.empty()) dst = k1;
                else
                    for (int q = 0; q < g.N[d]; ++q) dst[q] += k1[q];
            }
            j.has_T = true;
        }
    }
    if (las) {
        if ((long)las->size() < nsteps) fail(QDX_ERR_PARAM, "Clock values doesn't fit the laser field");
        j.laser = *las;
    }
    return j;
}

// Double-buffered device tables: the set written for a new job mix is not
// the one the ticks already enqueued read.
struct TableSet {
    DevBuf jobs, r256, r512, col, plans;
    hipEvent_t last_use = nullptr;
    bool used = false;
};

}  // namespace

struct qdx_batch {
    qdx_ctx *ctx = nullptr;
    int precision = QDX_FP64;
    Arena arena;
    std::map<int64_t, BatchJob> jobs;
    int64_t next_id = 1;
    int64_t ticks = 0;
    DevBuf dtick;
    TableSet sets[2];
    int cur = 0;
    bool dirty = true;
    // current tables
    std::vector<int64_t> order_ids;        // job index -> id
    int n256 = 0, n512 = 0, ncol = 0;
    size_t lds256 = 0, lds512 = 0, ldsc = 0;
    bool tw256 = true, tw512 = true, twc = true;
    size_t rsz() const { return precision == QDX_FP32_BULK ? sizeof(float) : sizeof(double); }
};

namespace {

template <class R>
void build_tables(qdx_batch *b) {
    const bool fp32 = std::is_same<R, float>::value;
    const size_t esz = 2 * sizeof(R);
    std::vector<BJob<R>> tj;
    std::vector<BTile> r256, r512, cols;
    std::vector<LinePlan> plans;
    std::map<int, int> pidx;
    auto plan_of = [&](int L) {
        auto it = pidx.find(L);
        if (it != pidx.end()) return it->second;
        plans.push_back(b->ctx->plan(L, fp32));
        pidx[L] = (int)plans.size() - 1;
        return (int)plans.size() - 1;
    };
    b->order_ids.clear();
    size_t l256 = 0, l256n = 0, l512 = 0, l512n = 0, lc = 0, lcn = 0;
    char *base = b->arena.base();
    for (auto &kv : b->jobs) {
        const BatchJob &J = kv.second;
        const int ji = (int)tj.size();
        b->order_ids.push_back(J.id);
        BJob<R> d{};
        char *jb = base + J.off;
        d.nd = J.grid.nd;
        d.N0 = J.grid.N[0];
        d.N1 = J.grid.nd == 2 ? J.grid.N[1] : 1;
        d.plan0 = plan_of(d.N0);
        d.plan1 = J.grid.nd == 2 ? plan_of(d.N1) : 0;
        d.rb = J.tile.rb;
        d.nrt = J.tile.nrt;
        d.W = J.tile.W;
        d.order = J.cp.order;
        d.start_tick = (int)J.start_tick;
        d.nsteps = J.nsteps;
        d.nt = J.nt;
        d.size = J.grid.size();
        d.inv_n = 1.0 / double(d.size);
        d.offset = J.cp.offset;
        d.inv_scaling = 1.0 / J.cp.scaling;
        d.prec = J.cp.prec;
        d.lscale = J.lscale;
        d.coef = reinterpret_cast<const double *>(jb + J.lay.coef);
        d.tk0 = J.has_T ? reinterpret_cast<const double *>(jb + J.lay.tk0) : nullptr;
        d.tk1 = J.has_T && J.grid.nd == 2 ? reinterpret_cast<const double *>(jb + J.lay.tk1) : nullptr;
        d.V = J.has_V ? reinterpret_cast<const R *>(jb + J.lay.V) : nullptr;
        d.mu = J.has_mu ? reinterpret_cast<const R *>(jb + J.lay.mu) : nullptr;
        d.laser = J.nt > 0 ? reinterpret_cast<const double *>(jb + J.lay.laser) : nullptr;
        d.psi = reinterpret_cast<double *>(jb + J.lay.psi);
        d.phiA = reinterpret_cast<R *>(jb + J.lay.A);
        d.phiB = reinterpret_cast<R *>(jb + J.lay.B);
        d.X = J.grid.nd == 2 ? reinterpret_cast<R *>(jb + J.lay.X) : nullptr;
        d.part = reinterpret_cast<double *>(jb + J.lay.part);
        d.ctl = reinterpret_cast<ChebyCtl *>(jb + J.lay.ctl);
        d.stats = reinterpret_cast<long long *>(jb + J.lay.stats);
        std::vector<BTile> &rl = J.tile.NT == 256 ? r256 : r512;
        d.rt0 = (int)rl.size();
        for (int r = 0; r < J.tile.nrt; ++r) rl.push_back(BTile{ji, r * J.tile.rb});
        const size_t lt = row_lds(J.grid, J.tile, esz, true), ln = row_lds(J.grid, J.tile, esz, false);
        if (J.tile.NT == 256) {
            l256 = std::max(l256, lt);
            l256n = std::max(l256n, ln);
        } else {
            l512 = std::max(l512, lt);
            l512n = std::max(l512n, ln);
        }
        if (J.grid.nd == 2) {
            for (int c = 0; c < J.tile.nct; ++c) cols.push_back(BTile{ji, c * J.tile.W});
            lc = std::max(lc, col_lds(J.grid, J.tile, esz, true));
            lcn = std::max(lcn, col_lds(J.grid, J.tile, esz, false));
        }
        tj.push_back(d);
    }
    b->tw256 = l256 <= kLds;
    b->lds256 = b->tw256 ? l256 : l256n;
    b->tw512 = l512 <= kLds;
    b->lds512 = b->tw512 ? l512 : l512n;
    b->twc = lc <= kLds;
    b->ldsc = b->twc ? lc : lcn;
    if (b->lds256 > kLds || b->lds512 > kLds || b->ldsc > kLds) fail(QDX_ERR_PARAM, "job tile exceeds the LDS");
    b->n256 = (int)r256.size();
    b->n512 = (int)r512.size();
    b->ncol = (int)cols.size();

    // write the tables into the set the enqueued ticks do not read
    const int nxt = 1 - b->cur;
    TableSet &ts = b->sets[nxt];
    if (ts.used) HIPCHK(hipEventSynchronize(ts.last_use));
    hipStream_t s = b->ctx->stream;
    auto put = [&](DevBuf &dst, const void *src, size_t bytes) {
        if (bytes == 0) return;
        if (dst.bytes < bytes) dst.alloc(bytes * 2);
        HIPCHK(hipMemcpyAsync(dst.p, src, bytes, hipMemcpyHostToDevice, s));
    };
    put(ts.jobs, tj.data(), tj.size() * sizeof(BJob<R>));
    put(ts.r256, r256.data(), r256.size() * sizeof(BTile));
    put(ts.r512, r512.data(), r512.size() * sizeof(BTile));
    put(ts.col, cols.data(), cols.size() * sizeof(BTile));
    put(ts.plans, plans.data(), plans.size() * sizeof(LinePlan));
    // the copies read pageable host memory: wait before the vectors go away
    HIPCHK(hipStreamSynchronize(s));
    b->cur = nxt;
    b->dirty = false;
}

template <class R>
void enqueue_tick(qdx_batch *b) {
    if (b->dirty) build_tables<R>(b);
    const int64_t T = b->ticks;
    bool any256 = false, any512 = false, any2d = false;
    int kmax = 0;
    for (auto &kv : b->jobs) {
        const BatchJob &J = kv.second;
        if (!J.active_at(T)) continue;
        (J.tile.NT == 256 ? any256 : any512) = true;
        if (J.grid.nd == 2) any2d = true;
        kmax = std::max(kmax, J.cp.order);
    }
    TableSet &ts = b->sets[b->cur];
    hipStream_t s = b->ctx->stream;
    const BJob<R> *jobs = ts.jobs.as<BJob<R>>();
    const LinePlan *plans = ts.plans.as<LinePlan>();
    const int *tick = b->dtick.as<int>();
    qdx_ctx *ctx = b->ctx;
    auto rows = [&](int phase, int k) {
        if (any256) {
            KTimer t(ctx, QDX_K_BATCH_ROW);
            BatchLaunch<R>::row(s, 256, phase, k, jobs, ts.r256.as<BTile>(), b->n256, plans, tick, b->lds256,
                                b->tw256);
        }
        if (any512) {
            KTimer t(ctx, QDX_K_BATCH_ROW);
            BatchLaunch<R>::row(s, 512, phase, k, jobs, ts.r512.as<BTile>(), b->n512, plans, tick, b->lds512,
                                b->tw512);
        }
    };
    auto cols = [&](int k) {
        if (!any2d) return;
        KTimer t(ctx, QDX_K_BATCH_COL);
        BatchLaunch<R>::col(s, k, jobs, ts.col.as<BTile>(), b->ncol, plans, tick, b->ldsc, b->twc);
    };
    if (kmax > 0) {
        rows(0, 0);
        cols(1);
        rows(1, 1);
        for (int k = 2; k < kmax; ++k) {
            cols(k);
            rows(2, k);
        }
    }
    BatchLaunch<R>::tick_end(s, jobs, (int)b->order_ids.size(), b->dtick.as<int>());
    if (!ts.last_use) HIPCHK(hipEventCreateWithFlags(&ts.last_use, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ts.last_use, s));
    ts.used = true;
    b->ticks++;
}

BatchJob &find_job(qdx_batch *b, int64_t id) {
    auto it = b->jobs.find(id);
    if (it == b->jobs.end()) fail(QDX_ERR_PARAM, "unknown job id");
    return it->second;
}

void check_desc(const qdx_job_desc *d) {
    if (!d || !d->ham || !d->psi0) fail(QDX_ERR_PARAM, "NULL argument");
    qdx_ham *h = d->ham;
    if (h->S != 1 || h->multistate) fail(QDX_ERR_INCOMPAT, "batched jobs hold one state");
    if (h->grid.nd > 2) fail(QDX_ERR_INCOMPAT, "batched jobs are 1D or 2D");
    if (h->grid.nd == 2 && h->grid.N[1] > 2048) fail(QDX_ERR_PARAM, "2D jobs need N1 <= 2048");
    if (d->nsteps < 1) fail(QDX_ERR_PARAM, "nsteps must be >= 1");
    if (d->dt == 0) fail(QDX_ERR_PARAM, "No time step defined");
}

// job parameters without touching the device
BatchJob plan_job(qdx_batch *b, const qdx_job_desc *d, JobHost &host) {
    check_desc(d);
    qdx_ham *h = d->ham;
    BatchJob J;
    J.grid = h->grid;
    J.nsteps = d->nsteps;
    host = job_host(h, d->nsteps);
    J.cp = cheby_params(h, d->dt, d->order, d->prec, 0.0);
    if (b->precision == QDX_FP32_BULK) fp32_series(J.cp, d->dt);
    J.tile = job_tiling(J.grid);
    J.has_V = !host.V.empty();
    J.has_mu = !host.mu.empty();
    J.has_T = host.has_T;
    J.lscale = host.lscale;
    J.nt = J.has_mu ? (int)host.laser.size() : 0;
    J.lay = job_layout(J.grid, J.tile, J.cp.order, J.nt, b->rsz(), J.has_mu);
    return J;
}

template <class R>
void upload_job(qdx_batch *b, BatchJob &J, const JobHost &host, const double *psi0, double dt) {
    char *jb = b->arena.base() + J.off;
    const long n = J.grid.size();
    auto put = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) HIPCHK(hipMemcpy(jb + off, src, bytes, hipMemcpyHostToDevice));
    };
    auto putR = [&](size_t off, const std::vector<double> &v) {
        if constexpr (std::is_same<R, float>::value) {
            std::vector<float> f(v.begin(), v.end());
            put(off, f.data(), f.size() * sizeof(float));
        } else {
            put(off, v.data(), v.size() * sizeof(double));
        }
    };
    put(J.lay.psi, psi0, 16 * (size_t)n);
    if (J.has_V) putR(J.lay.V, host.V);
    if (J.has_mu) {
        putR(J.lay.mu, host.mu);
        put(J.lay.laser, host.laser.data(), host.laser.size() * sizeof(double));
    }
    if (J.has_T) {
        put(J.lay.tk0, host.tk0.data(), host.tk0.size() * sizeof(double));
        if (J.grid.nd == 2) put(J.lay.tk1, host.tk1.data(), host.tk1.size() * sizeof(double));
    }
    std::vector<double> c(2 * J.cp.order);
    const std::complex<double> ph = std::exp(std::complex<double>(0, -1) * dt * J.cp.offset);
    for (int k = 0; k < J.cp.order; ++k) {
        const std::complex<double> v = J.cp.coeff[k] * ph;
        c[2 * k] = v.real();
        c[2 * k + 1] = v.imag();
    }
    put(J.lay.coef, c.data(), c.size() * sizeof(double));
    HIPCHK(hipMemset(jb + J.lay.ctl, 0, sizeof(ChebyCtl)));
    HIPCHK(hipMemset(jb + J.lay.stats, 0, 2 * sizeof(long long)));
}

}  // namespace

extern "C" {

int qdx_batch_create(qdx_ctx *ctx, int32_t precision, int64_t hbm_budget, qdx_batch **out) {
    return guarded([&] {
        if (!ctx || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (precision != QDX_FP64 && precision != QDX_FP32_BULK) fail(QDX_ERR_PARAM, "unknown precision");
        if (hbm_budget <= 0) fail(QDX_ERR_PARAM, "HBM budget must be positive");
        HIPCHK(hipSetDevice(ctx->device));
        auto b = std::make_unique<qdx_batch>();
        b->ctx = ctx;
        b->precision = precision;
        b->arena.init((size_t)hbm_budget);
        b->dtick.alloc(sizeof(int));
        HIPCHK(hipMemset(b->dtick.p, 0, sizeof(int)));
        HIPCHK(hipDeviceSynchronize());
        ctx_acquire(ctx);
        *out = b.release();
    });
}

int qdx_batch_footprint(qdx_batch *b, const qdx_job_desc *job, int64_t *bytes) {
    return guarded([&] {
        if (!b || !bytes) fail(QDX_ERR_PARAM, "NULL argument");
        JobHost host;
        const BatchJob J = plan_job(b, job, host);
        *bytes = (int64_t)J.lay.bytes;
    });
}

int qdx_batch_submit(qdx_batch *b, const qdx_job_desc *job, int64_t *id) {
    return guarded([&] {
        if (!b || !id) fail(QDX_ERR_PARAM, "NULL argument");
        JobHost host;
        BatchJob J = plan_job(b, job, host);
        if (!b->arena.alloc(J.lay.bytes, J.off))
            fail(QDX_ERR_MEMORY, "job does not fit into the free part of the batch's HBM budget");
        try {
            if (b->precision == QDX_FP32_BULK) upload_job<float>(b, J, host, job->psi0, job->dt);
            else upload_job<double>(b, J, host, job->psi0, job->dt);
        } catch (...) {
            b->arena.release(J.off, J.lay.bytes);
            throw;
        }
        J.id = b->next_id++;
        J.start_tick = b->ticks;
        b->jobs[J.id] = J;
        b->dirty = true;
        *id = J.id;
    });
}

int qdx_batch_run(qdx_batch *b, int32_t nticks) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        if (nticks < 0) fail(QDX_ERR_PARAM, "nticks must be >= 0");
        HIPCHK(hipSetDevice(b->ctx->device));
        for (int i = 0; i < nticks; ++i) {
            if (b->precision == QDX_FP32_BULK) enqueue_tick<float>(b);
            else enqueue_tick<double>(b);
        }
    });
}

int qdx_batch_job_state(qdx_batch *b, int64_t id, int64_t *steps_left, int64_t *n_hpsi, int64_t *steps_done) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        BatchJob &J = find_job(b, id);
        if (steps_left) *steps_left = std::max<int64_t>(0, J.start_tick + J.nsteps - b->ticks);
        if (n_hpsi || steps_done) {
            long long st[2];
            HIPCHK(hipMemcpyAsync(st, b->arena.base() + J.off + J.lay.stats, sizeof(st), hipMemcpyDeviceToHost,
                                  b->ctx->stream));
            HIPCHK(hipStreamSynchronize(b->ctx->stream));
            if (n_hpsi) *n_hpsi = st[0];
            if (steps_done) *steps_done = st[1];
        }
    });
}

int qdx_batch_fetch(qdx_batch *b, int64_t id, double *psi) {
    return guarded([&] {
        if (!b || !psi) fail(QDX_ERR_PARAM, "NULL argument");
        BatchJob &J = find_job(b, id);
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        HIPCHK(hipMemcpy(psi, b->arena.base() + J.off + J.lay.psi, 16 * (size_t)J.grid.size(),
                         hipMemcpyDeviceToHost));
    });
}

int qdx_batch_release(qdx_batch *b, int64_t id) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        BatchJob &J = find_job(b, id);
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        b->arena.release(J.off, J.lay.bytes);
        b->jobs.erase(id);
        b->dirty = true;
    });
}

int qdx_batch_usage(qdx_batch *b, int32_t *njobs, int64_t *bytes_used, int64_t *ticks) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        if (njobs) *njobs = (int32_t)b->jobs.size();
        if (bytes_used) *bytes_used = (int64_t)b->arena.used;
        if (ticks) *ticks = b->ticks;
    });
}

int qdx_batch_destroy(qdx_batch *b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        for (auto &ts : b->sets)
            if (ts.last_use) (void)hipEventDestroy(ts.last_use);
        qdx_ctx *ctx = b->ctx;
        delete b;
        ctx_release(ctx);
    });
}

}  // extern "C"
