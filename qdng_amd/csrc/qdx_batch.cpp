// Host runtime of the batched-jobs path (include/qdx.h, "batched jobs"):
// device arena, job admission by footprint, the job/tile tables the batched
// kernels read, the per-tick launch sequence and the continuous-batching
// serve loop.
#include "qdx_batch.hpp"
#include "qdx_host_internal.hpp"

#include <chrono>
#include <deque>
#include <thread>
#include <tuple>

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

// Sub-buffers of one job inside its arena block (offsets; the job's flags say
// which exist).
struct JobLayout {
    size_t psi, A, B, X, V, mu, V11, V01, tk0, tk1, tk0s1, tk1s1, laser, coef, part, ctl, stats, bytes;
};

// Tiling of one job: row tiles of rb lines (NT threads), column tiles of W
// columns (256 threads).
struct JobTiling {
    int NT = 256, rb = 1, nrt = 1, W = 1, nct = 0;
};

enum JobKind { K_CHEBY = 0, K_GSPO = 1 };

struct BatchJob {
    int64_t id = 0;
    int kind = K_CHEBY;
    int S = 1;
    Grid grid;
    ChebyParams cp;          // Chebyshev jobs
    double dt = 0;
    int nsteps = 0;
    int64_t start_tick = 0;
    size_t off = 0;
    JobLayout lay{};
    JobTiling tile;
    bool has_V = false, has_mu = false, has_T = false, has_V11 = false, has_V01 = false, has_T1 = false;
    std::vector<int> rad0, rad1;   // radix plans of N0 / N1, fixed at submission
    double lscale = 1.0;
    int nt = 0;
    int64_t last_tick() const { return start_tick + nsteps - 1; }
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

// Host arrays of one job.
struct JobHost {
    std::vector<double> V, mu, V11, V01, tk0, tk1, tk0s1, tk1s1, laser;
    double lscale = 1.0;
    bool has_T = false, has_T1 = false;
};

void add_kinetic(JobHost &j, const Grid &g, const Term &t) {
    for (int d = 0; d < g.nd; ++d) {
        std::vector<double> &dst = t.row == 0 ? (d == 0 ? j.tk0 : j.tk1) : (d == 0 ? j.tk0s1 : j.tk1s1);
        const std::vector<double> k1 = kspace_d2(g, d, t.mass[d]);
        if (dst.empty()) dst = k1;
        else
            for (int q = 0; q < g.N[d]; ++q) dst[q] += k1[q];
    }
    (t.row == 0 ? j.has_T : j.has_T1) = true;
}

void add_field(std::vector<double> &dst, const std::vector<double> &v) {
    if (dst.empty()) dst.assign(v.size(), 0.0);
    for (size_t x = 0; x < v.size(); ++x) dst[x] += v[x];
}

// Chebyshev job: H = OSum of grid terms on one state
JobHost cheby_host(qdx_ham *h, int nsteps, int first) {
    JobHost j;
    const std::vector<double> *las = nullptr;
    for (const Term &t : h->terms) {
        if (t.row != 0 || t.col != 0) fail(QDX_ERR_INCOMPAT, "batched Chebyshev jobs hold one state");
        if (t.kind == T_NIP) fail(QDX_ERR_INCOMPAT, "batched jobs take no negative imaginary potential");
        if (t.kind == T_POT) {
            add_field(j.V, t.v);
        } else if (t.kind == T_DIP) {
            if (las && (*las != t.laser || j.lscale != t.scale))
                fail(QDX_ERR_INCOMPAT, "two dipole terms with different fields on one element");
            add_field(j.mu, t.v);
            las = &t.laser;
            j.lscale = t.scale;
        } else {
            add_kinetic(j, h->grid, t);
        }
    }
    if (las) {   // the samples from the job's first clock step on
        if ((long)las->size() < (long)first + nsteps) fail(QDX_ERR_PARAM, "Clock values doesn't fit the laser field");
        j.laser.assign(las->begin() + first, las->end());
    }
    return j;
}

// GSPO job: operators {T, W} (OGSPO with A = kinetic, B = potential)
JobHost gspo_host(qdx_ham *T, qdx_ham *W, int nsteps, int first) {
    JobHost j;
    const int S = T->S;
    for (const Term &t : T->terms) {
        if (t.kind != T_KIN) fail(QDX_ERR_INCOMPAT, "the first GSPO operator must hold kinetic terms only");
        add_kinetic(j, T->grid, t);
    }
    const std::vector<double> *las = nullptr;
    bool c01 = false, c10 = false;
    for (const Term &t : W->terms) {
        if (t.kind == T_KIN) fail(QDX_ERR_INCOMPAT, "the second GSPO operator must hold potentials only");
        if (t.kind == T_NIP) fail(QDX_ERR_INCOMPAT, "batched jobs take no negative imaginary potential");
        if (t.kind == T_DIP) {
            if (S != 1) fail(QDX_ERR_INCOMPAT, "batched GSPO jobs take a dipole on a single state only");
            if (las && (*las != t.laser || j.lscale != t.scale))
                fail(QDX_ERR_INCOMPAT, "two dipole terms with different fields on one element");
            add_field(j.mu, t.v);
            las = &t.laser;
            j.lscale = t.scale;
            continue;
        }
        if (t.row == 0 && t.col == 0) add_field(j.V, t.v);
        else if (t.row == 1 && t.col == 1) add_field(j.V11, t.v);
        else {
            (t.row == 0 ? c01 : c10) = true;
            add_field(j.V01, t.v);
        }
    }
    if (c01 && c10) fail(QDX_ERR_INCOMPAT, "batched GSPO jobs take one coupling element (mirrored)");
    if (las) {   // the samples from the job's first clock step on
        if ((long)las->size() < (long)first + nsteps) fail(QDX_ERR_PARAM, "Clock values doesn't fit the laser field");
        j.laser.assign(las->begin() + first, las->end());
    }
    return j;
}

// rsz: bytes of a real operator value; tsz: bytes of one complex value of the
// Chebyshev term buffers (A, B, X; 10 in the compact 40-bit storage)
JobLayout job_layout(const BatchJob &J, size_t rsz, size_t tsz) {
    JobLayout l{};
    size_t o = 0;
    auto take = [&](size_t n) {
        const size_t at = o;
        o += Arena::round(std::max<size_t>(n, 1));
        return at;
    };
    const Grid &g = J.grid;
    const size_t n = (size_t)g.size(), S = (size_t)J.S;
    if (J.kind == K_CHEBY) {
        l.psi = take(16 * n);
        l.A = take(tsz * n);
        l.B = take(tsz * n);
        l.coef = take(16 * (size_t)J.cp.order);
        // room for the partials of one-line row tiles, so any tiling fits
        l.part = take(8 * 3 * (size_t)std::max(J.tile.nrt, J.grid.nd == 2 ? J.grid.N[1] : 1));
    } else {
        l.psi = take(2 * rsz * n * S);
    }
    if (g.nd == 2) l.X = take((J.kind == K_CHEBY ? tsz : 2 * rsz) * n * S);
    if (J.has_V) l.V = take(rsz * n);
    if (J.has_mu) l.mu = take(rsz * n);
    if (J.has_V11) l.V11 = take(rsz * n);
    if (J.has_V01) l.V01 = take(rsz * n);
    l.tk0 = take(8 * (size_t)g.N[0]);
    if (g.nd == 2) l.tk1 = take(8 * (size_t)g.N[1]);
    if (J.has_T1) {
        l.tk0s1 = take(8 * (size_t)g.N[0]);
        if (g.nd == 2) l.tk1s1 = take(8 * (size_t)g.N[1]);
    }
    if (J.has_mu) l.laser = take(8 * (size_t)std::max(J.nt, 1));
    l.ctl = take(sizeof(ChebyCtl));
    l.stats = take(2 * sizeof(long long));
    l.bytes = o;
    return l;
}

// LDS of the tiles: padded line buffers (L + L/8 entries per line) and the
// padded twiddle copy (L + L/8 entries) when it is staged
size_t lp(int L) { return (size_t)L + L / 8; }
// lb: LDS buffers per line (2 ping-pong, 1 in place)
size_t row_lds(const Grid &g, const JobTiling &t, size_t esz, bool tw, int lb = 2) {
    const int L = g.N[0];
    return (lb * t.rb * lp(L) + (tw ? lp(L) : 0)) * esz;
}
size_t col_lds(const Grid &g, const JobTiling &t, size_t esz, bool tw, int lb = 2) {
    const int L = g.N[1];
    return (lb * t.W * lp(L) + (tw ? lp(L) : 0)) * esz;
}
size_t line_lds(const Grid &g, size_t esz, bool tw) {
    const int L = g.N[0];
    return (2 * lp(L) + (tw ? lp(L) : 0)) * esz;
}

constexpr size_t kLds = 160 * 1024;

// LDS bytes of one launch: the largest tile, with the twiddle table staged in
// LDS when that still fits (else all blocks read the table from memory)
struct LdsNeed {
    size_t with = 0, without = 0;
    void add(size_t w, size_t wo) {
        with = std::max(with, w);
        without = std::max(without, wo);
    }
    bool tw() const { return with <= kLds; }
    size_t bytes() const { return tw() ? with : without; }
};

// One launch list: tiles (or job indices) with its LDS need and, for the
// per-order Chebyshev lists, (order, tile count) prefix marks.
struct TileList {
    std::vector<BTile> tiles;
    std::vector<int> jobs;
    LdsNeed lds;
    std::vector<std::pair<int, int>> pre;
    int n2d = 0;   // GSPO row lists: the 2D jobs' tiles come first
};

enum ListId { L_ROW256, L_ROW512, L_COL, L_LINE256, L_LINE512, L_GROW256, L_GROW512, L_GCOL, L_COUNT };

bool is_line_list(int i) { return i == L_LINE256 || i == L_LINE512; }

// Device tables, double-buffered: the set written for a new job mix is not
// the one the ticks already enqueued read.  Uploads go through pinned staging
// memory, so a rebuild does not wait for the device.
struct TableSet {
    DevBuf dev;
    char *host = nullptr;
    size_t cap = 0;
    size_t off_jobs = 0, off_plans = 0, off_list[L_COUNT] = {};
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
    int njobs_tab = 0;
    int count[L_COUNT] = {};
    size_t lds[L_COUNT] = {};
    bool twl[L_COUNT] = {};
    std::vector<std::pair<int, int>> pre[L_COUNT];
    int n2d[L_COUNT] = {};
    // 1D Chebyshev jobs run on a side stream that forks from and joins the
    // main stream every tick
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // Tick graphs.  The launch sequence of a tick depends only on the table
    // set (its generation `gen`), which job classes are active and the longest
    // active series; the kernels take the tick from the device counter.  Key:
    // (gen, class flags, kmax).  The first tick of a key runs eagerly (it also
    // raises the kernels' LDS limits), the second is captured and replayed.
    struct GraphSlot {
        int seen = 0;
        int set = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::map<std::tuple<uint64_t, int, int>, GraphSlot> graphs;
    uint64_t gen = 0;
    bool use_graphs = true;
    int64_t n_captured = 0, n_replayed = 0;
    // tilings of 2D Chebyshev job shapes (N0, N1), set or tuned
    std::map<std::pair<int, int>, JobTiling> tilings;
    bool autotune = false;
    // LDS buffers per line of the Chebyshev row / column tiles of all 2D jobs
    // (2: ping-pong, 1: in place; bit-identical results).  Default: in place,
    // since the LDS of two buffers, not the registers, bounds the waves per
    // SIMD of both kernels in both precisions (FP64 rows and columns 4
    // waves in place, 2 with two buffers; FP32 6 against 4).
    int lbr = 1, lbc = 1;
    size_t rsz() const { return precision == QDX_FP32_BULK ? sizeof(float) : sizeof(double); }
    // QDX_COMPACT: the terms of 2D Chebyshev jobs in the 40-bit storage
    bool cmp() const { return precision == QDX_COMPACT; }
};

namespace {

template <class R>
void build_tables(qdx_batch *b) {
    const bool fp32 = std::is_same<R, float>::value;
    const size_t esz = 2 * sizeof(R);
    std::vector<BJob<R>> tj;
    TileList L[L_COUNT];
    std::vector<LinePlan> plans;
    std::map<std::pair<int, std::vector<int>>, int> pidx;
    auto plan_of = [&](int len, const std::vector<int> &rad) {
        auto it = pidx.find({len, rad});
        if (it != pidx.end()) return it->second;
        plans.push_back(b->ctx->plan(len, fp32, rad));
        pidx[{len, rad}] = (int)plans.size() - 1;
        return (int)plans.size() - 1;
    };
    // Chebyshev jobs by decreasing series length (prefix launches per order),
    // GSPO 2D jobs before GSPO 1D jobs
    std::vector<const BatchJob *> order;
    for (auto &kv : b->jobs) order.push_back(&kv.second);
    std::stable_sort(order.begin(), order.end(), [](const BatchJob *x, const BatchJob *y) {
        if (x->kind != y->kind) return x->kind < y->kind;
        if (x->kind == K_CHEBY) return x->cp.order > y->cp.order;
        return x->grid.nd > y->grid.nd;
    });
    char *base = b->arena.base();
    for (const BatchJob *jp : order) {
        const BatchJob &J = *jp;
        const int ji = (int)tj.size();
        BJob<R> d{};
        char *jb = base + J.off;
        auto at = [&](size_t off) { return jb + off; };
        d.kind = J.kind;
        d.S = J.S;
        d.nd = J.grid.nd;
        d.N0 = J.grid.N[0];
        d.N1 = J.grid.nd == 2 ? J.grid.N[1] : 1;
        d.plan0 = plan_of(d.N0, J.rad0);
        d.plan1 = J.grid.nd == 2 ? plan_of(d.N1, J.rad1) : 0;
        d.rb = J.tile.rb;
        d.nrt = J.tile.nrt;
        d.W = J.tile.W;
        d.order = J.kind == K_CHEBY ? J.cp.order : 0;
        d.start_tick = (int)J.start_tick;
        d.nsteps = J.nsteps;
        d.nt = J.nt;
        d.size = J.grid.size();
        d.inv_n = 1.0 / double(d.size);
        d.inv_n0 = 1.0 / double(d.N0);
        d.inv_n1 = 1.0 / double(d.N1);
        d.tau = J.dt;
        d.lscale = J.lscale;
        d.tk0 = J.has_T ? reinterpret_cast<const double *>(at(J.lay.tk0)) : nullptr;
        d.tk1 = J.has_T && J.grid.nd == 2 ? reinterpret_cast<const double *>(at(J.lay.tk1)) : nullptr;
        d.tk0s1 = J.has_T1 ? reinterpret_cast<const double *>(at(J.lay.tk0s1)) : nullptr;
        d.tk1s1 = J.has_T1 && J.grid.nd == 2 ? reinterpret_cast<const double *>(at(J.lay.tk1s1)) : nullptr;
        d.V = J.has_V ? reinterpret_cast<const R *>(at(J.lay.V)) : nullptr;
        d.mu = J.has_mu ? reinterpret_cast<const R *>(at(J.lay.mu)) : nullptr;
        d.V11 = J.has_V11 ? reinterpret_cast<const R *>(at(J.lay.V11)) : nullptr;
        d.V01 = J.has_V01 ? reinterpret_cast<const R *>(at(J.lay.V01)) : nullptr;
        d.laser = J.has_mu ? reinterpret_cast<const double *>(at(J.lay.laser)) : nullptr;
        d.X = J.grid.nd == 2 ? reinterpret_cast<R *>(at(J.lay.X)) : nullptr;
        d.ctl = reinterpret_cast<ChebyCtl *>(at(J.lay.ctl));
        d.stats = reinterpret_cast<long long *>(at(J.lay.stats));
        if (J.kind == K_CHEBY) {
            d.offset = J.cp.offset;
            d.inv_scaling = 1.0 / J.cp.scaling;
            d.prec = J.cp.prec;
            d.coef = reinterpret_cast<const double *>(at(J.lay.coef));
            d.psi = reinterpret_cast<double *>(at(J.lay.psi));
            d.phiA = reinterpret_cast<R *>(at(J.lay.A));
            d.phiB = reinterpret_cast<R *>(at(J.lay.B));
            d.part = reinterpret_cast<double *>(at(J.lay.part));
            if (J.grid.nd == 1) {
                TileList &l = L[J.tile.NT == 256 ? L_LINE256 : L_LINE512];
                l.jobs.push_back(ji);
                l.lds.add(line_lds(J.grid, esz, true), line_lds(J.grid, esz, false));
            } else {
                TileList &l = L[J.tile.NT == 256 ? L_ROW256 : L_ROW512];
                d.rt0 = (int)l.tiles.size();
                for (int r = 0; r < J.tile.nrt; ++r) l.tiles.push_back(BTile{ji, r * J.tile.rb});
                l.pre.push_back({J.cp.order, (int)l.tiles.size()});
                l.lds.add(row_lds(J.grid, J.tile, esz, true, b->lbr), row_lds(J.grid, J.tile, esz, false, b->lbr));
                TileList &c = L[L_COL];
                for (int k = 0; k < J.tile.nct; ++k) c.tiles.push_back(BTile{ji, k * J.tile.W});
                c.pre.push_back({J.cp.order, (int)c.tiles.size()});
                c.lds.add(col_lds(J.grid, J.tile, esz, true, b->lbc), col_lds(J.grid, J.tile, esz, false, b->lbc));
            }
        } else {
            d.psiR = reinterpret_cast<R *>(at(J.lay.psi));
            TileList &l = L[J.tile.NT == 256 ? L_GROW256 : L_GROW512];
            d.rt0 = (int)l.tiles.size();
            for (int r = 0; r < J.tile.nrt; ++r) l.tiles.push_back(BTile{ji, r * J.tile.rb});
            if (J.grid.nd == 2) l.n2d = (int)l.tiles.size();
            l.lds.add(row_lds(J.grid, J.tile, esz, true), row_lds(J.grid, J.tile, esz, false));
            if (J.grid.nd == 2) {
                TileList &c = L[L_GCOL];
                for (int s = 0; s < J.S; ++s)
                    for (int k = 0; k < J.tile.nct; ++k) c.tiles.push_back(BTile{ji, s * d.N0 + k * J.tile.W});
                c.lds.add(col_lds(J.grid, J.tile, esz, true), col_lds(J.grid, J.tile, esz, false));
            }
        }
        tj.push_back(d);
    }
    for (int i = 0; i < L_COUNT; ++i) {
        if (L[i].lds.bytes() > kLds) fail(QDX_ERR_PARAM, "job tile exceeds the LDS");
        b->count[i] = (int)(is_line_list(i) ? L[i].jobs.size() : L[i].tiles.size());
        b->lds[i] = L[i].lds.bytes();
        b->twl[i] = L[i].lds.tw();
        b->pre[i] = L[i].pre;
        b->n2d[i] = L[i].n2d;
    }
    b->njobs_tab = (int)tj.size();

    // stage into pinned memory, one async copy into the set not in use
    const int nxt = 1 - b->cur;
    TableSet &ts = b->sets[nxt];
    if (ts.used) HIPCHK(hipEventSynchronize(ts.last_use));
    // graphs that read set nxt have completed (its last use has passed); the
    // uncaptured keys of older mixes are not looked up again
    for (auto it = b->graphs.begin(); it != b->graphs.end();) {
        if (it->second.exec && it->second.set != nxt) {
            ++it;
            continue;
        }
        if (it->second.exec) HIPCHK(hipGraphExecDestroy(it->second.exec));
        it = b->graphs.erase(it);
    }
    size_t o = 0;
    auto slot = [&](size_t bytes) {
        const size_t at = o;
        o += Arena::round(std::max<size_t>(bytes, 1));
        return at;
    };
    ts.off_jobs = slot(tj.size() * sizeof(BJob<R>));
    ts.off_plans = slot(plans.size() * sizeof(LinePlan));
    for (int i = 0; i < L_COUNT; ++i)
        ts.off_list[i] = slot(is_line_list(i) ? L[i].jobs.size() * sizeof(int) : L[i].tiles.size() * sizeof(BTile));
    if (o > ts.cap) {
        if (ts.host) HIPCHK(hipHostFree(ts.host));
        ts.host = nullptr;
        ts.cap = 0;
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&ts.host), 2 * o, hipHostMallocDefault));
        ts.cap = 2 * o;
        ts.dev.alloc(2 * o);
    }
    auto put = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) std::memcpy(ts.host + off, src, bytes);
    };
    put(ts.off_jobs, tj.data(), tj.size() * sizeof(BJob<R>));
    put(ts.off_plans, plans.data(), plans.size() * sizeof(LinePlan));
    for (int i = 0; i < L_COUNT; ++i) {
        if (is_line_list(i)) put(ts.off_list[i], L[i].jobs.data(), L[i].jobs.size() * sizeof(int));
        else put(ts.off_list[i], L[i].tiles.data(), L[i].tiles.size() * sizeof(BTile));
    }
    HIPCHK(hipMemcpyAsync(ts.dev.p, ts.host, o, hipMemcpyHostToDevice, b->ctx->stream));
    b->cur = nxt;
    b->dirty = false;
    b->gen++;
}

// tiles of the Chebyshev jobs whose series reaches order k
int tiles_for(const std::vector<std::pair<int, int>> &pre, int k) {
    int n = 0;
    for (const auto &e : pre) {
        if (e.first <= k) break;
        n = e.second;
    }
    return n;
}

template <class R>
void enqueue_tick(qdx_batch *b) {
    if (b->dirty) build_tables<R>(b);
    const int64_t T = b->ticks;
    bool cheb2d = false, cheb1d = false, gspo = false;
    int kmax = 0;
    for (auto &kv : b->jobs) {
        const BatchJob &J = kv.second;
        if (!J.active_at(T)) continue;
        if (J.kind == K_GSPO) {
            gspo = true;
        } else if (J.grid.nd == 1) {
            cheb1d = true;
        } else {
            cheb2d = true;
            kmax = std::max(kmax, J.cp.order);
        }
    }
    TableSet &ts = b->sets[b->cur];
    hipStream_t s = b->ctx->stream;
    char *tb = ts.dev.as<char>();
    const BJob<R> *jobs = reinterpret_cast<const BJob<R> *>(tb + ts.off_jobs);
    const LinePlan *plans = reinterpret_cast<const LinePlan *>(tb + ts.off_plans);
    auto tiles = [&](int i) { return reinterpret_cast<const BTile *>(tb + ts.off_list[i]); };
    auto list = [&](int i) { return reinterpret_cast<const int *>(tb + ts.off_list[i]); };
    const int *tick = b->dtick.as<int>();
    qdx_ctx *ctx = b->ctx;
    auto body = [&] {
        if (cheb1d) {
            // the 1D jobs' steps overlap the other jobs' passes
            HIPCHK(hipEventRecord(b->fork, s));
            HIPCHK(hipStreamWaitEvent(b->side, b->fork, 0));
            for (int i : {L_LINE256, L_LINE512}) {
                if (b->count[i] == 0) continue;
                KTimer t(ctx, QDX_K_BATCH_LINE);
                BatchLaunch<R>::line(b->side, i == L_LINE256 ? 256 : 512, jobs, list(i), b->count[i], plans, tick,
                                     b->lds[i], b->twl[i]);
            }
            HIPCHK(hipEventRecord(b->join, b->side));
        }
        if (gspo) {
            for (int i : {L_GROW256, L_GROW512}) {
                if (b->count[i] == 0) continue;
                KTimer t(ctx, QDX_K_BATCH_ROW);
                BatchLaunch<R>::gspo_row(s, i == L_GROW256 ? 256 : 512, 0, jobs, tiles(i), b->count[i], plans, tick,
                                         b->lds[i], b->twl[i]);
            }
            if (b->count[L_GCOL] > 0) {
                KTimer t(ctx, QDX_K_BATCH_COL);
                BatchLaunch<R>::gspo_col(s, jobs, tiles(L_GCOL), b->count[L_GCOL], plans, tick, b->lds[L_GCOL],
                                         b->twl[L_GCOL]);
            }
            for (int i : {L_GROW256, L_GROW512}) {
                if (b->n2d[i] == 0) continue;
                KTimer t(ctx, QDX_K_BATCH_ROW);
                BatchLaunch<R>::gspo_row(s, i == L_GROW256 ? 256 : 512, 1, jobs, tiles(i), b->n2d[i], plans, tick,
                                         b->lds[i], b->twl[i]);
            }
        }
        auto rows = [&](int phase, int k) {
            for (int i : {L_ROW256, L_ROW512}) {
                const int n = tiles_for(b->pre[i], k);
                if (n == 0) continue;
                KTimer t(ctx, QDX_K_BATCH_ROW);
                BatchLaunch<R>::row(s, i == L_ROW256 ? 256 : 512, b->lbr, phase, k, jobs, tiles(i), n, plans, tick,
                                    b->lds[i], b->twl[i], b->cmp());
            }
        };
        auto cols = [&](int k) {
            const int n = tiles_for(b->pre[L_COL], k);
            if (n == 0) return;
            KTimer t(ctx, QDX_K_BATCH_COL);
            BatchLaunch<R>::col(s, b->lbc, k, jobs, tiles(L_COL), n, plans, tick, b->lds[L_COL], b->twl[L_COL],
                                b->cmp());
        };
        if (cheb2d) {
            rows(0, 0);
            cols(1);
            rows(1, 1);
            for (int k = 2; k < kmax; ++k) {
                cols(k);
                rows(2, k);
            }
        }
        if (cheb1d) HIPCHK(hipStreamWaitEvent(s, b->join, 0));
        BatchLaunch<R>::tick_end(s, jobs, b->njobs_tab, b->dtick.as<int>());
    };
    if (!b->use_graphs || ctx->profiling) {
        body();   // launches carry the profiling events
    } else {
        const int flags = (cheb2d ? 1 : 0) | (cheb1d ? 2 : 0) | (gspo ? 4 : 0);
        auto &g = b->graphs[std::make_tuple(b->gen, flags, kmax)];
        g.set = b->cur;
        if (g.exec) {
            HIPCHK(hipGraphLaunch(g.exec, s));
            b->n_replayed++;
        } else if (g.seen++ == 0) {
            body();
        } else {
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                hipGraph_t dead = nullptr;
                (void)hipStreamEndCapture(s, &dead);
                if (dead) (void)hipGraphDestroy(dead);
                throw;
            }
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamEndCapture(s, &graph));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(e);
            g.exec = exec;
            b->n_captured++;
            HIPCHK(hipGraphLaunch(exec, s));
            b->n_replayed++;
        }
    }
    if (!ts.last_use) HIPCHK(hipEventCreateWithFlags(&ts.last_use, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ts.last_use, s));
    ts.used = true;
    b->ticks++;
}

void tick(qdx_batch *b) {
    if (b->precision == QDX_FP32_BULK) enqueue_tick<float>(b);
    else enqueue_tick<double>(b);
}

BatchJob &find_job(qdx_batch *b, int64_t id) {
    auto it = b->jobs.find(id);
    if (it == b->jobs.end()) fail(QDX_ERR_PARAM, "unknown job id");
    return it->second;
}

void check_grid(const Grid &g) {
    if (g.nd > 2) fail(QDX_ERR_INCOMPAT, "batched jobs are 1D or 2D");
    // the batched transforms (in-place schedule) run radix 8/4/2/3 stages only
    for (int d = 0; d < g.nd; ++d)
        if (!is_smooth23(g.N[d])) fail(QDX_ERR_INCOMPAT, "batched jobs take line lengths 8 * 2^a * 3^b");
}

// job parameters and host arrays, without touching the device
BatchJob plan_job(qdx_batch *b, const qdx_job_desc *d, JobHost &host) {
    if (!d || !d->psi0) fail(QDX_ERR_PARAM, "NULL argument");
    if (d->nsteps < 1) fail(QDX_ERR_PARAM, "nsteps must be >= 1");
    if (d->first_step < 0) fail(QDX_ERR_PARAM, "first_step must be >= 0");
    if (d->dt == 0) fail(QDX_ERR_PARAM, "No time step defined");
    BatchJob J;
    J.nsteps = d->nsteps;
    J.dt = d->dt;
    if (d->nops == 0) {
        qdx_ham *h = d->ham;
        if (!h) fail(QDX_ERR_PARAM, "NULL argument");
        if (h->S != 1 || h->multistate) fail(QDX_ERR_INCOMPAT, "batched Chebyshev jobs hold one state");
        J.kind = K_CHEBY;
        J.S = 1;
        J.grid = h->grid;
        check_grid(J.grid);
        host = cheby_host(h, d->nsteps, d->first_step);
        J.cp = cheby_params(h, d->dt, d->order, d->prec, 0.0);
        // fixed-length series for FP32 terms and for the compact terms of 2D
        // jobs (1D jobs of a compact batch keep their terms in registers and
        // run as FP64 jobs)
        if (b->precision == QDX_FP32_BULK || (b->cmp() && J.grid.nd == 2)) fp32_series(J.cp, d->dt);
        J.cp.order = launch_window(J.cp, d->dt);   // orders launched (and coefficients uploaded)
    } else {
        if (d->nops != 2 || !d->ops || !d->ops[0] || !d->ops[1])
            fail(QDX_ERR_INCOMPAT, "batched GSPO jobs take the split operators {T, W}");
        qdx_ham *T = d->ops[0], *W = d->ops[1];
        if (!(T->grid == W->grid) || T->S != W->S) fail(QDX_ERR_INCOMPAT, "GSPO operators on different grids");
        if (T->S > 2) fail(QDX_ERR_INCOMPAT, "batched GSPO jobs hold one or two states");
        J.kind = K_GSPO;
        J.S = T->S;
        J.grid = T->grid;
        check_grid(J.grid);
        host = gspo_host(T, W, d->nsteps, d->first_step);
    }
    const bool fp32 = b->precision == QDX_FP32_BULK;
    J.rad0 = b->ctx->radix_of(J.grid.N[0], fp32);
    if (J.grid.nd == 2) J.rad1 = b->ctx->radix_of(J.grid.N[1], fp32);
    J.tile = job_tiling(J.grid);
    if (J.kind == K_CHEBY && J.grid.nd == 2) {
        auto it = b->tilings.find({J.grid.N[0], J.grid.N[1]});
        if (it != b->tilings.end()) J.tile = it->second;
    }
    J.has_V = !host.V.empty();
    J.has_mu = !host.mu.empty();
    J.has_V11 = !host.V11.empty();
    J.has_V01 = !host.V01.empty();
    J.has_T = host.has_T;
    J.has_T1 = host.has_T1;
    J.lscale = host.lscale;
    J.nt = J.has_mu ? (int)host.laser.size() : 0;
    const bool cmp_terms = b->cmp() && J.kind == K_CHEBY && J.grid.nd == 2;
    J.lay = job_layout(J, b->rsz(), cmp_terms ? 10 : 2 * b->rsz());
    return J;
}

template <class R>
void upload_job(qdx_batch *b, BatchJob &J, const JobHost &host, const double *psi0) {
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
    if (J.kind == K_CHEBY) {
        put(J.lay.psi, psi0, 16 * (size_t)n);
        std::vector<double> c(2 * J.cp.order);
        const std::complex<double> ph = std::exp(std::complex<double>(0, -1) * J.dt * J.cp.offset);
        for (int k = 0; k < J.cp.order; ++k) {
            const std::complex<double> v = J.cp.coeff[k] * ph;
            c[2 * k] = v.real();
            c[2 * k + 1] = v.imag();
        }
        put(J.lay.coef, c.data(), c.size() * sizeof(double));
    } else {
        putR(J.lay.psi, std::vector<double>(psi0, psi0 + 2 * (size_t)n * J.S));
    }
    if (J.has_V) putR(J.lay.V, host.V);
    if (J.has_V11) putR(J.lay.V11, host.V11);
    if (J.has_V01) putR(J.lay.V01, host.V01);
    if (J.has_mu) {
        putR(J.lay.mu, host.mu);
        put(J.lay.laser, host.laser.data(), host.laser.size() * sizeof(double));
    }
    if (J.has_T) {
        put(J.lay.tk0, host.tk0.data(), host.tk0.size() * sizeof(double));
        if (J.grid.nd == 2) put(J.lay.tk1, host.tk1.data(), host.tk1.size() * sizeof(double));
    }
    if (J.has_T1) {
        put(J.lay.tk0s1, host.tk0s1.data(), host.tk0s1.size() * sizeof(double));
        if (J.grid.nd == 2) put(J.lay.tk1s1, host.tk1s1.data(), host.tk1s1.size() * sizeof(double));
    }
    const ChebyCtl c0{};
    put(J.lay.ctl, &c0, sizeof(c0));
    const long long z[2] = {0, 0};
    put(J.lay.stats, z, sizeof(z));
}

void tune_shape(qdx_batch *b, int n0, int n1, qdx_tune_result *res);

// admit a job; false when it does not fit into the free part of the arena
bool try_submit(qdx_batch *b, const qdx_job_desc *d, int64_t &id) {
    if (b->autotune && d && d->nops == 0 && d->ham && d->ham->grid.nd == 2 &&
        !b->tilings.count({d->ham->grid.N[0], d->ham->grid.N[1]}))
        tune_shape(b, d->ham->grid.N[0], d->ham->grid.N[1], nullptr);
    JobHost host;
    BatchJob J = plan_job(b, d, host);
    if (!b->arena.alloc(J.lay.bytes, J.off)) return false;
    try {
        if (b->precision == QDX_FP32_BULK) upload_job<float>(b, J, host, d->psi0);
        else upload_job<double>(b, J, host, d->psi0);
    } catch (...) {
        b->arena.release(J.off, J.lay.bytes);
        throw;
    }
    J.id = b->next_id++;
    J.start_tick = b->ticks;
    id = J.id;
    b->jobs[id] = std::move(J);
    b->dirty = true;
    return true;
}

// psi of a job whose ticks have completed (complex128)
void copy_psi(qdx_batch *b, const BatchJob &J, double *out) {
    const char *src = b->arena.base() + J.off + J.lay.psi;
    const size_t n = (size_t)J.grid.size() * J.S;
    if (J.kind == K_CHEBY || b->precision != QDX_FP32_BULK) {
        HIPCHK(hipMemcpy(out, src, 16 * n, hipMemcpyDeviceToHost));
        return;
    }
    std::vector<float> f(2 * n);
    HIPCHK(hipMemcpy(f.data(), src, f.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 2 * n; ++i) out[i] = f[i];
}

void read_stats(qdx_batch *b, const BatchJob &J, long long st[2]) {
    HIPCHK(hipMemcpy(st, b->arena.base() + J.off + J.lay.stats, 2 * sizeof(long long), hipMemcpyDeviceToHost));
}

// -1, or the step whose state's norm was not a number (the job stopped there)
int64_t failed_step(qdx_batch *b, const BatchJob &J) {
    ChebyCtl c;
    HIPCHK(hipMemcpy(&c, b->arena.base() + J.off + J.lay.ctl, sizeof(c), hipMemcpyDeviceToHost));
    return c.failed > 0 ? (int64_t)c.failed - 1 : -1;
}

// return a job's memory; the caller has made sure no enqueued tick still uses it
void drop_job(qdx_batch *b, int64_t id) {
    BatchJob &J = find_job(b, id);
    b->arena.release(J.off, J.lay.bytes);
    b->jobs.erase(id);
    b->dirty = true;
}

bool any_steps_left(const qdx_batch *b) {
    for (auto &kv : b->jobs)
        if (kv.second.active_at(b->ticks)) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Schedules and the autotuner

Grid shape_grid(int n0, int n1) {
    Grid g;
    g.nd = 2;
    g.N[0] = n0;
    g.N[1] = n1;
    check_grid(g);
    radix_plan(n0);
    radix_plan(n1);
    if (n0 > kMaxLine || n1 > kMaxLine) fail(QDX_ERR_PARAM, "batched lines hold at most 4096 points");
    return g;
}

// a tiling from its public description, checked against the kernels' limits
JobTiling tiling_from(const Grid &g, const qdx_tiling &q, size_t esz) {
    const int TL = g.N[0] / 8, TL1 = g.N[1] / 8;
    if (q.row_threads != 256 && q.row_threads != 512) fail(QDX_ERR_PARAM, "row tiles use 256 or 512 threads");
    if (q.row_lines < 1 || q.row_lines * TL > q.row_threads || q.row_lines > g.N[1])
        fail(QDX_ERR_PARAM, "row_lines * N0/8 must not exceed row_threads (and row_lines <= N1)");
    if (q.col_width < 1 || q.col_width * TL1 > 256 || g.N[0] % q.col_width != 0)
        fail(QDX_ERR_PARAM, "col_width * N1/8 must not exceed 256 and col_width must divide N0");
    JobTiling t;
    t.NT = q.row_threads;
    t.rb = q.row_lines;
    t.nrt = (g.N[1] + t.rb - 1) / t.rb;
    t.W = q.col_width;
    t.nct = g.N[0] / t.W;
    if (row_lds(g, t, esz, false) > kLds || col_lds(g, t, esz, false) > kLds)
        fail(QDX_ERR_PARAM, "tile exceeds the 160 KiB LDS of a gfx950 CU");
    return t;
}

qdx_tiling tiling_to(const JobTiling &t) {
    qdx_tiling q;
    q.row_threads = t.NT;
    q.row_lines = t.rb;
    q.col_width = t.W;
    return q;
}

void rc_check(int rc) {
    if (rc != QDX_OK) fail(rc, qdx_last_error());
}

// K synthetic 2D jobs of one shape in a context and batch of their own, for
// timing: kinetic term (mass 2000 per dimension), harmonic potential, dipole
// x exp(-x^2/8) in a sine field, Gaussian psi0, a fixed 16-term series
// without the break test (every tick runs the same orders), about 2 M grid
// points in all.
struct TuneBench {
    static constexpr int kTicks = 4, kSteps = 1000;
    qdx_ctx *tc = nullptr;
    qdx_batch *tb = nullptr;
    qdx_ham *h = nullptr;
    int K = 1;
    TuneBench(int device, int precision, const Grid &g) {
        rc_check(qdx_ctx_create(device, &tc));
        qdx_grid q{};
        q.ndims = 2;
        for (int d = 0; d < 2; ++d) {
            q.N[d] = g.N[d];
            q.xmin[d] = -10.0;
            q.xmax[d] = 10.0;
        }
        rc_check(qdx_ham_create(nullptr, &q, 1, 0, 1, &h));
        const double mass[2] = {2000.0, 2000.0};
        rc_check(qdx_ham_add_kinetic(h, 0, mass));
        const long n = g.size();
        std::vector<double> V(n), mu(n), psi(2 * n), laser(kSteps);
        const double dx0 = 20.0 / (g.N[0] - 1), dx1 = 20.0 / (g.N[1] - 1);
        for (int j = 0; j < g.N[1]; ++j)
            for (int i = 0; i < g.N[0]; ++i) {
                const double x = -10.0 + i * dx0, y = -10.0 + j * dx1;
                const long p = (long)j * g.N[0] + i;
                V[p] = 0.02 * (x * x + y * y);
                mu[p] = x * std::exp(-x * x / 8);
                psi[2 * p] = std::exp(-(x - 1) * (x - 1) - y * y);
                psi[2 * p + 1] = 0.0;
            }
        for (int t = 0; t < kSteps; ++t) laser[t] = 0.01 * std::sin(0.05 * t);
        rc_check(qdx_ham_add_potential(h, 0, 0, V.data()));
        rc_check(qdx_ham_add_dipole(h, 0, 0, mu.data(), laser.data(), kSteps, 1.0));
        double emin = 0, emax = 0;
        rc_check(qdx_ham_bounds(h, &emin, &emax));
        qdx_job_desc d{};
        d.ham = h;
        d.nsteps = kSteps;
        d.dt = 8.0 / (0.5 * (emax - emin));   // R dt = 8 < 16 terms
        d.order = 16;
        d.prec = 0.0;
        d.psi0 = psi.data();
        K = (int)std::max<long>(1, std::min<long>(512, (2L << 20) / n));
        int64_t fp = 0;
        rc_check(qdx_batch_create(tc, precision, 1 << 20, &tb));
        rc_check(qdx_batch_footprint(tb, &d, &fp));
        rc_check(qdx_batch_destroy(tb));
        tb = nullptr;
        rc_check(qdx_batch_create(tc, precision, (int64_t)K * (fp + 4096), &tb));
        for (int k = 0; k < K; ++k) {
            int64_t id = 0;
            rc_check(qdx_batch_submit(tb, &d, &id));
        }
    }
    ~TuneBench() {
        if (tb) qdx_batch_destroy(tb);
        if (h) qdx_ham_destroy(h);
        if (tc) qdx_ctx_destroy(tc);
    }
    void set(const std::vector<int> &r0, const std::vector<int> &r1, const JobTiling &t) {
        for (auto &kv : tb->jobs) {
            kv.second.rad0 = r0;
            kv.second.rad1 = r1;
            kv.second.tile = t;
        }
        tb->dirty = true;
    }
    // row and column kernel time per tick (us) of the current schedule
    std::pair<double, double> measure() {
        rc_check(qdx_batch_run(tb, 1));   // tables, kernel LDS limits, caches
        rc_check(qdx_ctx_synchronize(tc));
        rc_check(qdx_ctx_profile(tc, 1));
        rc_check(qdx_batch_run(tb, kTicks));
        int64_t n = 0;
        double row = 0, col = 0;
        rc_check(qdx_ctx_kernel_time(tc, QDX_K_BATCH_ROW, &n, &row));
        rc_check(qdx_ctx_kernel_time(tc, QDX_K_BATCH_COL, &n, &col));
        rc_check(qdx_ctx_profile(tc, 0));
        return {row * 1e3 / kTicks, col * 1e3 / kTicks};
    }
};

// radix plans worth timing for length L: plans with no more stages than
// the default one, fewest stages first (at most 12), and four with one
// stage more
std::vector<std::vector<int>> plan_choices(int L) {
    std::vector<std::vector<int>> out, more;
    const size_t nmin = radix_plan(L).size();
    std::vector<std::vector<int>> all = radix_candidates(L, 1);
    std::stable_sort(all.begin(), all.end(),
                     [](const std::vector<int> &a, const std::vector<int> &b) { return a.size() < b.size(); });
    for (auto &r : all) {
        if (r.size() <= nmin) {
            if (out.size() < 12) out.push_back(r);
        } else if (more.size() < 4) {
            more.push_back(r);
        }
    }
    out.insert(out.end(), more.begin(), more.end());
    return out;
}

void tune_shape(qdx_batch *b, int n0, int n1, qdx_tune_result *res) {
    const Grid g = shape_grid(n0, n1);
    const size_t esz = 2 * b->rsz();
    TuneBench tb(b->ctx->device, b->precision, g);
    tb.tb->lbr = b->lbr;   // the batch's LDS schedule
    tb.tb->lbc = b->lbc;
    // a candidate replaces the best only when it is 2 % faster (timing noise)
    auto better = [](double t, double best) { return t < 0.98 * best; };
    std::vector<int> r0 = radix_plan(n0), r1 = radix_plan(n1);
    JobTiling bt = job_tiling(g);
    int ncand = 1;
    tb.set(r0, r1, bt);
    const auto base = tb.measure();
    double brow = base.first, bcol = base.second;
    // radix plans (a square grid shares one plan between rows and columns)
    if (n0 == n1) {
        for (auto &c : plan_choices(n0)) {
            if (c == r0) continue;
            tb.set(c, c, bt);
            const auto m = tb.measure();
            ++ncand;
            if (better(m.first + m.second, brow + bcol)) {
                r0 = r1 = c;
                brow = m.first;
                bcol = m.second;
            }
        }
    } else {
        for (auto &c : plan_choices(n0)) {
            if (c == r0) continue;
            tb.set(c, r1, bt);
            const auto m = tb.measure();
            ++ncand;
            if (better(m.first, brow)) {
                r0 = c;
                brow = m.first;
            }
        }
        for (auto &c : plan_choices(n1)) {
            if (c == r1) continue;
            tb.set(r0, c, bt);
            const auto m = tb.measure();
            ++ncand;
            if (better(m.second, bcol)) {
                r1 = c;
                bcol = m.second;
            }
        }
    }
    // Row tiles stay at the default block size: a different one would put
    // this shape's tiles into a launch of their own (row lists are split by
    // block size), and in a mix of shapes the extra launch per order costs
    // more than a shape gains alone (measured: 64 x 64 FP32 rows 7 % faster
    // alone in 512-thread blocks, the mixed tick slower).  Full blocks of the
    // default size leave no other choice of lines per tile.
    const int TL1 = n1 / 8;
    // column tiles: width 1 .. 16
    const JobTiling col0 = bt;
    for (int W = 1; W <= 16; W *= 2) {
        if (W * TL1 > 256 || n0 % W != 0) continue;
        JobTiling t = col0;
        t.W = W;
        t.nct = n0 / W;
        if (t.W == col0.W) continue;
        if (col_lds(g, t, esz, false) > kLds) continue;
        tb.set(r0, r1, t);
        const auto m = tb.measure();
        ++ncand;
        if (better(m.second, bcol)) {
            bt.W = t.W;
            bt.nct = t.nct;
            bcol = m.second;
        }
    }
    const bool fp32 = b->precision == QDX_FP32_BULK;
    b->ctx->radix[{n0, fp32}] = r0;
    b->ctx->radix[{n1, fp32}] = r1;
    b->tilings[{n0, n1}] = bt;
    if (res) {
        *res = qdx_tune_result{};
        res->tiling = tiling_to(bt);
        res->nst0 = (int32_t)r0.size();
        res->nst1 = (int32_t)r1.size();
        for (size_t i = 0; i < r0.size(); ++i) res->radix0[i] = r0[i];
        for (size_t i = 0; i < r1.size(); ++i) res->radix1[i] = r1[i];
        res->row_us_default = base.first;
        res->col_us_default = base.second;
        res->row_us = brow;
        res->col_us = bcol;
        res->candidates = ncand;
    }
}

// LDS schedule of the current mix: every (row, column) buffer choice timed on
// the batch's own jobs (the choice does not change results, so the ticks are
// ordinary steps of the jobs).  Returns the ms per tick of (2,2), (1,2),
// (2,1), (1,1) in ms[4].
void tune_lds(qdx_batch *b, double ms[4]) {
    // every active job runs through all 12 ticks, so each schedule is timed
    // on the same work
    if (!any_steps_left(b)) fail(QDX_ERR_STATE, "no active jobs to time the LDS schedule on");
    for (auto &kv : b->jobs) {
        const BatchJob &J = kv.second;
        if (J.active_at(b->ticks) && J.start_tick + J.nsteps - b->ticks < 12)
            fail(QDX_ERR_STATE, "an active job has fewer than 12 steps left");
    }
    qdx_ctx *ctx = b->ctx;
    if (ctx->profiling) fail(QDX_ERR_STATE, "LDS schedule tuning needs the context's profiling off");
    const int modes[4][2] = {{2, 2}, {1, 2}, {2, 1}, {1, 1}};
    double best = 0;
    int bi = -1;
    // kernel by kernel: a graph capture would fall into the timed ticks
    struct GraphsOff {
        qdx_batch *b;
        bool was;
        ~GraphsOff() { b->use_graphs = was; }
    } off{b, b->use_graphs};
    b->use_graphs = false;
    for (int m = 0; m < 4; ++m) {
        b->lbr = modes[m][0];
        b->lbc = modes[m][1];
        b->dirty = true;
        tick(b);   // tables of this schedule
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < 2; ++k) tick(b);
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ms[m] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / 2;
        if (bi < 0 || ms[m] < best) {
            best = ms[m];
            bi = m;
        }
    }
    b->lbr = modes[bi][0];
    b->lbc = modes[bi][1];
    b->dirty = true;
}

}  // namespace

extern "C" {

int qdx_batch_create(qdx_ctx *ctx, int32_t precision, int64_t hbm_budget, qdx_batch **out) {
    return guarded([&] {
        if (!ctx || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (precision != QDX_FP64 && precision != QDX_FP32_BULK && precision != QDX_COMPACT)
            fail(QDX_ERR_PARAM, "unknown precision");
        if (hbm_budget <= 0) fail(QDX_ERR_PARAM, "HBM budget must be positive");
        HIPCHK(hipSetDevice(ctx->device));
        auto b = std::make_unique<qdx_batch>();
        b->ctx = ctx;
        b->precision = precision;
        b->lbr = 1;
        b->lbc = 1;
        b->arena.init((size_t)hbm_budget);
        b->dtick.alloc(sizeof(int));
        HIPCHK(hipMemset(b->dtick.p, 0, sizeof(int)));
        HIPCHK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
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
        *bytes = (int64_t)Arena::round(J.lay.bytes);
    });
}

int qdx_batch_submit(qdx_batch *b, const qdx_job_desc *job, int64_t *id) {
    return guarded([&] {
        if (!b || !id) fail(QDX_ERR_PARAM, "NULL argument");
        HIPCHK(hipSetDevice(b->ctx->device));
        if (!try_submit(b, job, *id))
            fail(QDX_ERR_MEMORY, "job does not fit into the free part of the batch's HBM budget");
    });
}

int qdx_batch_run(qdx_batch *b, int32_t nticks) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        if (nticks < 0) fail(QDX_ERR_PARAM, "nticks must be >= 0");
        HIPCHK(hipSetDevice(b->ctx->device));
        for (int i = 0; i < nticks; ++i) tick(b);
    });
}

int qdx_batch_job_state(qdx_batch *b, int64_t id, int64_t *steps_left, int64_t *n_hpsi, int64_t *steps_done) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        BatchJob &J = find_job(b, id);
        if (steps_left) *steps_left = std::max<int64_t>(0, J.start_tick + J.nsteps - b->ticks);
        if (n_hpsi || steps_done) {
            HIPCHK(hipStreamSynchronize(b->ctx->stream));
            long long st[2];
            read_stats(b, J, st);
            if (n_hpsi) *n_hpsi = st[0];
            if (steps_done) *steps_done = st[1];
        }
    });
}

int qdx_batch_job_status(qdx_batch *b, int64_t id, int64_t *failed_at) {
    return guarded([&] {
        if (!b || !failed_at) fail(QDX_ERR_PARAM, "NULL argument");
        BatchJob &J = find_job(b, id);
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        *failed_at = failed_step(b, J);
    });
}

int qdx_batch_fetch(qdx_batch *b, int64_t id, double *psi) {
    return guarded([&] {
        if (!b || !psi) fail(QDX_ERR_PARAM, "NULL argument");
        BatchJob &J = find_job(b, id);
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        copy_psi(b, J, psi);
    });
}

int qdx_batch_release(qdx_batch *b, int64_t id) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        find_job(b, id);
        HIPCHK(hipStreamSynchronize(b->ctx->stream));
        drop_job(b, id);
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

int qdx_batch_serve(qdx_batch *b, const qdx_job_desc *jobs, const double *arrival_s, int32_t n,
                    int32_t max_active, double *const *psi_out, qdx_serve_result *out) {
    std::vector<hipEvent_t> pool;
    hipEvent_t ev0 = nullptr;
    int rc = guarded([&] {
        if (!b || (n > 0 && (!jobs || !arrival_s || !out))) fail(QDX_ERR_PARAM, "NULL argument");
        for (int i = 1; i < n; ++i)
            if (arrival_s[i] < arrival_s[i - 1]) fail(QDX_ERR_PARAM, "arrival times must not decrease");
        HIPCHK(hipSetDevice(b->ctx->device));
        hipStream_t s = b->ctx->stream;
        using clk = std::chrono::steady_clock;
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(hipEventCreate(&ev0));
        const auto t0 = clk::now();
        HIPCHK(hipEventRecord(ev0, s));
        auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
        struct Flight {
            int64_t tick;
            hipEvent_t ev;
        };
        std::deque<Flight> flights;
        std::deque<int> pending;
        std::map<int64_t, int> active;              // job id -> request
        std::multimap<int64_t, int64_t> ending;     // last tick -> job id
        constexpr size_t kAhead = 4;                // ticks in flight before the host waits
        int next = 0, done = 0;
        while (done < n) {
            const double t = now();
            while (next < n && arrival_s[next] <= t) pending.push_back(next++);
            while (!pending.empty() && (max_active <= 0 || (int)active.size() < max_active)) {
                const int i = pending.front();
                int64_t id;
                if (!try_submit(b, &jobs[i], id)) break;
                out[i].admit_s = now();
                active[id] = i;
                ending.insert({b->jobs[id].last_tick(), id});
                pending.pop_front();
            }
            const bool work = any_steps_left(b);
            if (!work && flights.empty()) {
                if (!pending.empty() && active.empty())
                    fail(QDX_ERR_MEMORY, "a request does not fit into the whole HBM budget");
                if (next < n) {
                    const double wait = arrival_s[next] - now();
                    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
                }
                continue;
            }
            if (work) {
                tick(b);
                hipEvent_t ev;
                if (pool.empty()) {
                    HIPCHK(hipEventCreate(&ev));
                } else {
                    ev = pool.back();
                    pool.pop_back();
                }
                HIPCHK(hipEventRecord(ev, s));
                flights.push_back({b->ticks - 1, ev});
            }
            // retire the jobs whose last tick has completed; wait only when
            // too many ticks are in flight or nothing else can be enqueued
            while (!flights.empty()) {
                Flight &f = flights.front();
                const bool block = flights.size() > kAhead || !work;
                const hipError_t q = block ? hipEventSynchronize(f.ev) : hipEventQuery(f.ev);
                if (q == hipErrorNotReady) break;
                HIPCHK(q);
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev0, f.ev));
                auto range = ending.equal_range(f.tick);
                for (auto it = range.first; it != range.second; ++it) {
                    const int64_t id = it->second;
                    const int i = active[id];
                    const BatchJob &J = b->jobs[id];
                    long long st[2];
                    read_stats(b, J, st);
                    out[i].finish_s = ms * 1e-3;
                    out[i].n_hpsi = st[0];
                    out[i].failed_step = failed_step(b, J);
                    if (psi_out && psi_out[i]) copy_psi(b, J, psi_out[i]);
                    drop_job(b, id);
                    active.erase(id);
                    ++done;
                }
                ending.erase(range.first, range.second);
                pool.push_back(f.ev);
                flights.pop_front();
                if (block && flights.size() <= kAhead && work) break;
            }
        }
        HIPCHK(hipStreamSynchronize(s));
    });
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    return rc;
}

int qdx_batch_set_graphs(qdx_batch *b, int32_t enable) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        b->use_graphs = enable != 0;
    });
}

int qdx_batch_graph_stats(qdx_batch *b, int64_t *captured, int64_t *replayed) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        if (captured) *captured = b->n_captured;
        if (replayed) *replayed = b->n_replayed;
    });
}

int qdx_batch_get_tiling(qdx_batch *b, int32_t n0, int32_t n1, qdx_tiling *t) {
    return guarded([&] {
        if (!b || !t) fail(QDX_ERR_PARAM, "NULL argument");
        const Grid g = shape_grid(n0, n1);
        auto it = b->tilings.find({n0, n1});
        *t = tiling_to(it != b->tilings.end() ? it->second : job_tiling(g));
    });
}

int qdx_batch_set_tiling(qdx_batch *b, int32_t n0, int32_t n1, const qdx_tiling *t) {
    return guarded([&] {
        if (!b || !t) fail(QDX_ERR_PARAM, "NULL argument");
        const Grid g = shape_grid(n0, n1);
        b->tilings[{n0, n1}] = tiling_from(g, *t, 2 * b->rsz());
    });
}

int qdx_batch_autotune(qdx_batch *b, int32_t n0, int32_t n1, qdx_tune_result *out) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        HIPCHK(hipSetDevice(b->ctx->device));
        tune_shape(b, n0, n1, out);
    });
}

int qdx_batch_set_lds_schedule(qdx_batch *b, int32_t row_bufs, int32_t col_bufs) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        if ((row_bufs != 1 && row_bufs != 2) || (col_bufs != 1 && col_bufs != 2))
            fail(QDX_ERR_PARAM, "LDS buffers per line are 1 or 2");
        b->lbr = row_bufs;
        b->lbc = col_bufs;
        b->dirty = true;
    });
}

int qdx_batch_get_lds_schedule(qdx_batch *b, int32_t *row_bufs, int32_t *col_bufs) {
    return guarded([&] {
        if (!b || !row_bufs || !col_bufs) fail(QDX_ERR_PARAM, "NULL argument");
        *row_bufs = b->lbr;
        *col_bufs = b->lbc;
    });
}

int qdx_batch_autotune_lds(qdx_batch *b, double *ms_per_tick) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        HIPCHK(hipSetDevice(b->ctx->device));
        double ms[4] = {0, 0, 0, 0};
        tune_lds(b, ms);
        if (ms_per_tick)
            for (int i = 0; i < 4; ++i) ms_per_tick[i] = ms[i];
    });
}

int qdx_batch_set_autotune(qdx_batch *b, int32_t enable) {
    return guarded([&] {
        if (!b) fail(QDX_ERR_PARAM, "NULL argument");
        b->autotune = enable != 0;
    });
}

int qdx_batch_tuning_save(qdx_batch *b, const char *path) {
    return guarded([&] {
        if (!b || !path) fail(QDX_ERR_PARAM, "NULL argument");
        FILE *f = std::fopen(path, "w");
        if (!f) fail(QDX_ERR_PARAM, std::string("cannot write ") + path);
        const bool fp32 = b->precision == QDX_FP32_BULK;
        for (auto &kv : b->ctx->radix) {
            if (kv.first.second != fp32) continue;
            std::fprintf(f, "plan %d", kv.first.first);
            for (int r : kv.second) std::fprintf(f, " %d", r);
            std::fprintf(f, "\n");
        }
        for (auto &kv : b->tilings) {
            const JobTiling &t = kv.second;
            std::fprintf(f, "tile %d %d %d %d %d\n", kv.first.first, kv.first.second, t.NT, t.rb, t.W);
        }
        std::fprintf(f, "lds %d %d\n", b->lbr, b->lbc);
        const bool ok = std::fclose(f) == 0;
        if (!ok) fail(QDX_ERR_PARAM, std::string("cannot write ") + path);
    });
}

int qdx_batch_tuning_load(qdx_batch *b, const char *path) {
    return guarded([&] {
        if (!b || !path) fail(QDX_ERR_PARAM, "NULL argument");
        FILE *f = std::fopen(path, "r");
        if (!f) fail(QDX_ERR_PARAM, std::string("cannot read ") + path);
        std::map<int, std::vector<int>> plans;
        std::map<std::pair<int, int>, JobTiling> tiles;
        int lds_r = 0, lds_c = 0;
        char line[512];
        std::string bad;
        while (bad.empty() && std::fgets(line, sizeof line, f)) {
            char kind[8] = {};
            int used = 0;
            if (std::sscanf(line, "%7s%n", kind, &used) != 1) continue;   // blank line
            const char *rest = line + used;
            if (std::strcmp(kind, "plan") == 0) {
                int L = 0, r = 0, n = 0;
                std::vector<int> seq;
                if (std::sscanf(rest, "%d%n", &L, &n) != 1) bad = line;
                rest += n;
                while (bad.empty() && std::sscanf(rest, "%d%n", &r, &n) == 1) {
                    seq.push_back(r);
                    rest += n;
                }
                if (bad.empty() && (!is_smooth23(L) || !radix_ok(L, seq))) bad = line;
                plans[L] = seq;
            } else if (std::strcmp(kind, "lds") == 0) {
                if (std::sscanf(rest, "%d %d", &lds_r, &lds_c) != 2 || (lds_r != 1 && lds_r != 2) ||
                    (lds_c != 1 && lds_c != 2))
                    bad = line;
            } else if (std::strcmp(kind, "tile") == 0) {
                int n0, n1;
                qdx_tiling q;
                if (std::sscanf(rest, "%d %d %d %d %d", &n0, &n1, &q.row_threads, &q.row_lines, &q.col_width) != 5) {
                    bad = line;
                } else {
                    try {
                        tiles[{n0, n1}] = tiling_from(shape_grid(n0, n1), q, 2 * b->rsz());
                    } catch (const QdxError &) {
                        bad = line;
                    }
                }
            } else {
                bad = line;
            }
        }
        std::fclose(f);
        if (!bad.empty()) fail(QDX_ERR_PARAM, std::string("tuning file: bad line: ") + bad);
        for (auto &kv : plans) b->ctx->radix[{kv.first, b->precision == QDX_FP32_BULK}] = kv.second;
        for (auto &kv : tiles) b->tilings[kv.first] = kv.second;
        if (lds_r) {
            b->lbr = lds_r;
            b->lbc = lds_c;
            b->dirty = true;
        }
    });
}

int qdx_batch_destroy(qdx_batch *b) {
    return guarded([&] {
        if (!b) return;
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        (void)hipStreamSynchronize(b->side);
        for (auto &kv : b->graphs)
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        for (auto &ts : b->sets) {
            if (ts.last_use) (void)hipEventDestroy(ts.last_use);
            if (ts.host) (void)hipHostFree(ts.host);
        }
        if (b->fork) (void)hipEventDestroy(b->fork);
        if (b->join) (void)hipEventDestroy(b->join);
        if (b->side) (void)hipStreamDestroy(b->side);
        qdx_ctx *ctx = b->ctx;
        delete b;
        ctx_release(ctx);
    });
}

}  // extern "C"
