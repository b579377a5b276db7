// Optimal control (QDng ProgOCT, src/ProgOCT.cpp) over the device
// propagators: the Krotov, Rabitz and Rabitz-feedback iterations with
// overlap or projection targets, the laser samples of the Hamiltonian's
// dipole terms, and the projection operator.  Every wave-function operation
// is a call of the C ABI (device kernels); the field update of one step needs
// three scalar products on the host, so each step synchronises once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "qdx_host_internal.hpp"

struct qdx_proj {
    std::vector<qdx_wf *> wfs;   // copies, owned
    double sign = 1.0;
    bool inv = false;
    qdx_wf *buf = nullptr;       // scratch of in-place applications
    ~qdx_proj() {
        for (qdx_wf *w : wfs) qdx_wf_destroy(w);
        if (buf) qdx_wf_destroy(buf);
    }
};

namespace {

using cplx = std::complex<double>;

void chk(int rc) {
    if (rc != QDX_OK) fail(rc, g_err);
}

qdx_wf *new_like(const qdx_wf *like) {
    qdx_grid g{};
    g.ndims = like->grid.nd;
    for (int d = 0; d < like->grid.nd; ++d) {
        g.N[d] = like->grid.N[d];
        g.xmin[d] = like->grid.xmin[d];
        g.xmax[d] = like->grid.xmax[d];
    }
    qdx_wf *x = nullptr;
    chk(qdx_wf_create(like->ctx, &g, like->S, &x));
    return x;
}

struct WfSet {   // owns device wave functions
    std::vector<qdx_wf *> w;
    ~WfSet() {
        for (qdx_wf *x : w) qdx_wf_destroy(x);
    }
    qdx_wf *make(const qdx_wf *like) {
        w.push_back(nullptr);
        w.back() = new_like(like);
        return w.back();
    }
};

// <a|b>
cplx dot(qdx_wf *a, qdx_wf *b) {
    double o[2];
    chk(qdx_wf_overlap(a, b, o));
    return {o[0], o[1]};
}

double norm_of(qdx_wf *w) {
    double n = 0;
    chk(qdx_wf_norm(w, &n, nullptr));
    return n;
}

bool same_shape(const qdx_wf *a, const qdx_wf *b) { return a->ctx == b->ctx && a->grid == b->grid && a->S == b->S; }

// the Hamiltonians whose dipole terms carry the field: the propagator's own,
// or a GSPO propagator's split operators
std::vector<qdx_ham *> field_holders(qdx_prop *U) {
    std::vector<qdx_ham *> hs;
    if (U->kind == qdx_prop::GSPO) hs = U->ops;
    else if (U->ham) hs.push_back(U->ham);
    return hs;
}

std::vector<Term *> dipole_terms(qdx_ham *h) {
    std::vector<Term *> d;
    for (Term &t : h->terms)
        if (t.kind == T_DIP) d.push_back(&t);
    return d;
}

long laser_min_len(const std::vector<Term *> &dip) {
    long n = -1;
    for (const Term *t : dip) n = n < 0 ? (long)t->laser.size() : std::min(n, (long)t->laser.size());
    return n;
}

// ProgOCT: the Hamiltonian's lasers (_laserf), the previous field (_laserb),
// the coupling operator (_Coup) and the iteration state
struct Oct {
    qdx_prop *U = nullptr;
    qdx_oct_opts o{};
    const double *shape = nullptr;
    std::vector<qdx_wf *> psi_i, psi_t;
    std::vector<qdx_proj *> targets;
    qdx_ham *gobbler = nullptr;
    int nt = 1;                   // targets
    long steps = 0;
    std::vector<Term *> dip;      // every dipole term that carries the field
    double lscale = 1.0;          // Laser::Scale of the first laser
    double laser_dt = 0;
    std::vector<double> Ef, Eb;   // _laserf, _laserb samples
    double laserobj = 0;
    std::unique_ptr<qdx_ham, int (*)(qdx_ham *)> coup{nullptr, qdx_ham_destroy};
    WfSet own;
    std::vector<std::vector<qdx_wf *>> membuf;   // [target][0..steps]
    qdx_wf *opwf = nullptr, *gbuf = nullptr;

    // the field the Hamiltonian applies is Ef
    void push_field() {
        for (Term *t : dip) std::copy(Ef.begin(), Ef.end(), t->laser.begin());
    }
    void set_field(long s, double v) {   // Laser::Set on every laser
        Ef[s] = v;
        for (Term *t : dip) t->laser[s] = v;
    }
    void swap_fields() {   // Laser::swap of every laser pair
        std::swap(Ef, Eb);
        push_field();
    }
    void direction(bool backward) { chk(qdx_prop_set_direction(U, backward ? 1 : 0)); }

    void step(qdx_wf *w, long s) {
        chk(qdx_prop_apply(U, w, s, nullptr));
        if (gobbler) {   // _Gobbler->Apply(wf); wf->Normalize()
            chk(qdx_ham_apply(gobbler, gbuf, w, s));
            chk(qdx_wf_copy(w, gbuf));
            chk(qdx_wf_normalize(w));
        }
    }
    void store(int t, long idx, qdx_wf *w) { chk(qdx_wf_copy(membuf[t][idx], w)); }
    void copy_set(std::vector<qdx_wf *> &dst, const std::vector<qdx_wf *> &src) {
        for (int t = 0; t < nt; ++t) chk(qdx_wf_copy(dst[t], src[t]));
    }

    // ProgOCT::CalcCorr (ProgOCT.cpp:305-329)
    double corr(const std::vector<qdx_wf *> &wfi, const std::vector<qdx_wf *> &wft) {
        double res = 0;
        for (int t = 0; t < nt; ++t) {
            chk(qdx_ham_apply(coup.get(), opwf, wfi[t], 0));
            cplx im = o.phase ? dot(wft[t], opwf) : dot(wfi[t], wft[t]) * dot(wft[t], opwf);
            res += im.imag();
        }
        return res;
    }
    // ProgOCT::CalcLaserField (ProgOCT.cpp:334-361) at clock step s
    double field(const std::vector<qdx_wf *> &wfi, const std::vector<qdx_wf *> &wft, long s) {
        const double im = corr(wfi, wft);
        const double sh = shape[s];
        double res = 0;
        if (o.method == QDX_OCT_KROTOV) {
            res = Eb[s] * lscale - sh / (o.alpha * double(nt)) * im;
            const double a = sh / o.alpha * im;
            laserobj += a * a;
        } else {
            res = -1.0 * sh / (o.alpha * double(nt)) * im;
            laserobj += -1.0 / double(nt) * im * res;
        }
        if (std::isnan(res)) fail(QDX_ERR_OVERFLOW, "Laserfield is not a number");
        return res;
    }

    // ProgOCT::PropagateForward (ProgOCT.cpp:433-452): membuf[0] = t = 0
    void forward(std::vector<qdx_wf *> &wf) {
        for (int t = 0; t < nt; ++t) store(t, 0, wf[t]);
        for (long s = 0; s < steps; ++s)
            for (int t = 0; t < nt; ++t) {
                step(wf[t], s);
                store(t, s + 1, wf[t]);
            }
    }
    // ProgOCT::PropagateBackward (ProgOCT.cpp:460-492): T => 0 with the
    // previous field (the lasers exchanged around it)
    void backward(std::vector<qdx_wf *> &wf) {
        swap_fields();
        direction(true);
        for (int t = 0; t < nt; ++t) store(t, steps, wf[t]);
        for (long s = steps - 1; s >= 0; --s)
            for (int t = 0; t < nt; ++t) {
                step(wf[t], s);
                store(t, s, wf[t]);
            }
        swap_fields();
        direction(false);
    }
    void apply_target(int t, qdx_wf *dst, qdx_wf *src) { chk(qdx_proj_apply(targets[t], dst, src)); }

    // ProgOCT::Iterate (ProgOCT.cpp:541-581) with SyncTargetOverlap /
    // SyncTargetOperator (ProgOCT.cpp:500-535)
    void iterate(std::vector<qdx_wf *> &phii, std::vector<qdx_wf *> &phit) {
        if (!o.operator_target) {
            backward(phit);
        } else {
            swap_fields();
            forward(phit);
            swap_fields();
            for (int t = 0; t < nt; ++t) apply_target(t, phit[t], phit[t]);
            backward(phit);
        }
        copy_set(phii, psi_i);
        std::vector<qdx_wf *> bT(nt);
        for (long s = 0; s < steps; ++s) {
            for (int t = 0; t < nt; ++t) bT[t] = membuf[t][s];
            set_field(s, field(phii, bT, s));
            for (int t = 0; t < nt; ++t) step(phii[t], s);
        }
        for (int t = 0; t < nt; ++t) chk(qdx_wf_copy(phit[t], membuf[t][steps]));
    }

    // ProgOCT::IterateRabitzFB (ProgOCT.cpp:586-690)
    void iterate_fb(std::vector<qdx_wf *> &phii, std::vector<qdx_wf *> &phit, int it) {
        if (it == 1) forward(phii);
        for (int t = 0; t < nt; ++t) chk(qdx_wf_copy(phii[t], membuf[t][0]));
        if (o.operator_target)
            for (int t = 0; t < nt; ++t) apply_target(t, phit[t], phii[t]);
        std::vector<qdx_wf *> bT(nt);
        swap_fields();
        direction(true);
        for (long s = steps - 1; s >= 0; --s) {
            for (int t = 0; t < nt; ++t) bT[t] = membuf[t][s + 1];
            Eb[s] = field(bT, phit, s);   // _laserb->Set: not the applied field
            for (int t = 0; t < nt; ++t) store(t, s + 1, phit[t]);
            for (int t = 0; t < nt; ++t) step(phit[t], s);
        }
        for (int t = 0; t < nt; ++t) store(t, 0, phit[t]);
        swap_fields();
        direction(false);
        for (int t = 0; t < nt; ++t) chk(qdx_wf_copy(phit[t], membuf[t][steps]));
        for (long s = 0; s < steps; ++s) {
            for (int t = 0; t < nt; ++t) bT[t] = membuf[t][s];
            set_field(s, field(phii, bT, s));
            for (int t = 0; t < nt; ++t) store(t, s, phii[t]);
            for (int t = 0; t < nt; ++t) step(phii[t], s);
        }
        for (int t = 0; t < nt; ++t) store(t, steps, phii[t]);
    }

    // Laser::PulseEnergy (Laser.cpp:174-184) of the applied field
    double pulse_energy() const {
        double d = 0;
        for (double e : Ef) d += e * e;
        d *= lscale;
        return 137.0359996 / (4 * M_PI) * d * laser_dt;
    }

    // ProgOCT::Report (ProgOCT.cpp:363-425); returns |sum| and fills a row
    double report(const std::vector<qdx_wf *> &wfi, const std::vector<qdx_wf *> &wft, int it, double *row) {
        const double nan = std::nan("");
        cplx ov_sum = 0;
        double *pt = row ? row + QDX_OCT_ROW_FIXED : nullptr;
        for (int t = 0; t < nt; ++t) {
            double vals[QDX_OCT_ROW_PER_TARGET] = {nan, nan, nan, nan, nan};
            const double d = norm_of(wfi[t]);
            if (std::isnan(d)) fail(QDX_ERR_OVERFLOW, "Forward norm is not a number");
            vals[0] = d;
            if (!o.operator_target) {
                const double dn = norm_of(wft[t]);
                if (std::isnan(dn)) fail(QDX_ERR_OVERFLOW, "Backward norm is not a number");
                vals[1] = dn;
                const cplx ov = dot(wfi[t], wft[t]);
                vals[2] = std::abs(ov) * std::abs(ov);
                if (o.phase) {
                    vals[3] = std::arg(ov) / M_PI;
                    ov_sum += ov;
                } else {
                    ov_sum += vals[2];
                }
            } else {
                double e = 0;
                chk(qdx_proj_expec(targets[t], wfi[t], &e));
                vals[4] = e;
                ov_sum += std::abs(e);
            }
            if (pt) std::copy(vals, vals + QDX_OCT_ROW_PER_TARGET, pt + t * QDX_OCT_ROW_PER_TARGET);
        }
        const double y = std::abs(ov_sum);
        if (row) {
            row[0] = it;
            row[1] = it == 0 ? nan : laserobj;
            row[2] = it == 0 ? nan : ov_sum.real() - laserobj;
            row[3] = pulse_energy();
            row[4] = y;
        }
        return y;
    }
};

// the coupling operator of ProgOCT (FindOperatorType<OLaser>, ChainLoader.h:
// 48-160, with unity off, ProgOCT.cpp:954-961): a single-state Sum of the
// dipole arrays, or the multistate operator of the off-diagonal ones
// (mirrored like every lone off-diagonal element); arrays as OGridDipole
// holds them, without the field
qdx_ham *coupling_operator(qdx_prop *U, const std::vector<qdx_ham *> &hs) {
    qdx_ham *h0 = hs.front();
    const int S = h0->S;
    qdx_grid g{};
    g.ndims = h0->grid.nd;
    for (int d = 0; d < g.ndims; ++d) {
        g.N[d] = h0->grid.N[d];
        g.xmin[d] = h0->grid.xmin[d];
        g.xmax[d] = h0->grid.xmax[d];
    }
    qdx_ham *c = nullptr;
    chk(qdx_ham_create(U->ctx, &g, S, S > 1 ? 1 : 0, 0, &c));
    std::unique_ptr<qdx_ham, int (*)(qdx_ham *)> hold(c, qdx_ham_destroy);
    int added = 0;
    for (qdx_ham *h : hs)
        for (const Term &t : h->terms) {
            if (t.kind != T_DIP) continue;
            if (S > 1 && t.row == t.col) continue;   // diagonal dipoles do not couple
            chk(qdx_ham_add_potential(c, t.row, t.col, t.v.data()));
            ++added;
        }
    if (!added) fail(QDX_ERR_PARAM, "Given type of coupling operator not found in hamiltonian");
    return hold.release();
}

}  // namespace

extern "C" {

int qdx_ham_laser_len(qdx_ham *h, int64_t *n) {
    return guarded([&] {
        if (!h || !n) fail(QDX_ERR_PARAM, "NULL argument");
        *n = laser_min_len(dipole_terms(h));
        if (*n < 0) fail(QDX_ERR_PARAM, "the Hamiltonian has no dipole term");
    });
}

int qdx_ham_laser_get(qdx_ham *h, double *laser, int64_t n) {
    return guarded([&] {
        if (!h || !laser) fail(QDX_ERR_PARAM, "NULL argument");
        const std::vector<Term *> d = dipole_terms(h);
        if (d.empty()) fail(QDX_ERR_PARAM, "the Hamiltonian has no dipole term");
        if (n < 0 || n > (int64_t)d.front()->laser.size()) fail(QDX_ERR_PARAM, "n exceeds the laser length");
        std::copy(d.front()->laser.begin(), d.front()->laser.begin() + n, laser);
    });
}

int qdx_ham_laser_set(qdx_ham *h, const double *laser, int64_t n) {
    return guarded([&] {
        if (!h || !laser) fail(QDX_ERR_PARAM, "NULL argument");
        const std::vector<Term *> d = dipole_terms(h);
        if (d.empty()) fail(QDX_ERR_PARAM, "the Hamiltonian has no dipole term");
        if (n < 0 || n > laser_min_len(d)) fail(QDX_ERR_PARAM, "n exceeds the laser length");
        for (Term *t : d) std::copy(laser, laser + n, t->laser.begin());
    });
}

int qdx_proj_create(qdx_wf *const *wfs, int32_t n, int32_t positive, int32_t inv, qdx_proj **out) {
    return guarded([&] {
        if (!wfs || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (n < 1) fail(QDX_ERR_PARAM, "Projector: no wave functions");
        auto p = std::make_unique<qdx_proj>();
        for (int i = 0; i < n; ++i) {
            if (!wfs[i] || !wfs[i]->ctx) fail(QDX_ERR_PARAM, "Invalid WaveFunction");
            if (!same_shape(wfs[i], wfs[0])) fail(QDX_ERR_INCOMPAT, "Projector: wave functions differ in shape");
            p->wfs.push_back(nullptr);
            p->wfs.back() = new_like(wfs[i]);
            chk(qdx_wf_copy(p->wfs.back(), wfs[i]));
        }
        p->buf = new_like(wfs[0]);
        p->sign = positive ? 1.0 : -1.0;
        p->inv = inv != 0;
        *out = p.release();
    });
}

int qdx_proj_apply(qdx_proj *p, qdx_wf *dst, qdx_wf *src) {
    return guarded([&] {
        if (!p || !dst || !src) fail(QDX_ERR_PARAM, "NULL argument");
        if (!same_shape(dst, p->buf) || !same_shape(src, p->buf)) fail(QDX_ERR_INCOMPAT, "Projector: wrong shape");
        // OProjection::Apply(dest, src): dest = sign sum_i <wf_i|src> wf_i,
        // with inv src - that; built in a buffer when dst is src
        std::vector<cplx> c;
        for (qdx_wf *w : p->wfs) c.push_back(dot(w, src) * p->sign);
        qdx_wf *acc = dst == src ? p->buf : dst;
        double f = 1.0;
        if (p->inv) {
            chk(qdx_wf_copy(acc, src));
            f = -1.0;
        } else {
            HIPCHK(hipMemsetAsync(acc->ptr(), 0, (size_t)acc->S * acc->n() * 16, acc->ctx->stream));
        }
        for (size_t i = 0; i < p->wfs.size(); ++i) chk(qdx_wf_axpy(acc, f * c[i].real(), f * c[i].imag(), p->wfs[i]));
        if (acc != dst) chk(qdx_wf_copy(dst, acc));
    });
}

int qdx_proj_expec(qdx_proj *p, qdx_wf *psi, double *out) {
    return guarded([&] {
        if (!p || !psi || !out) fail(QDX_ERR_PARAM, "NULL argument");
        qdx_wf *tmp = new_like(psi);
        std::unique_ptr<qdx_wf, int (*)(qdx_wf *)> hold(tmp, qdx_wf_destroy);
        chk(qdx_proj_apply(p, tmp, psi));
        *out = dot(psi, tmp).real();
    });
}

int qdx_proj_destroy(qdx_proj *p) {
    delete p;
    return QDX_OK;
}

int qdx_oct_run(qdx_prop *U, const qdx_oct_opts *opts, const double *shape, int64_t nshape, qdx_wf *const *psi_i,
                qdx_wf *const *psi_t, qdx_proj *const *targets, qdx_ham *gobbler, double *report,
                int32_t *rows_done, int32_t *converged) {
    return guarded([&] {
        if (!U || !opts || !shape || !psi_i) fail(QDX_ERR_PARAM, "NULL argument");
        Oct c;
        c.U = U;
        c.o = *opts;
        c.shape = shape;
        c.gobbler = gobbler;
        c.nt = opts->ntargets;
        c.steps = opts->nsteps;
        // ProgOCT::_InitParams checks (ProgOCT.cpp:59-240)
        if (c.steps <= 0) fail(QDX_ERR_PARAM, "Zero number of time step defined");
        if (c.nt < 1) fail(QDX_ERR_PARAM, "Number of targets to small");
        if (c.nt > 16) fail(QDX_ERR_PARAM, "Number of targets exceeding MAX_TARGETS");
        if (opts->iterations < 1) fail(QDX_ERR_PARAM, "Invalid number of iterations");
        if (opts->conv < 0) fail(QDX_ERR_PARAM, "Invalid convergency criteria");
        if (std::abs(opts->alpha) == 0) fail(QDX_ERR_PARAM, "Invalid alpha parameter");
        if (opts->method < QDX_OCT_KROTOV || opts->method > QDX_OCT_RABITZFB)
            fail(QDX_ERR_PARAM, "Unknown OCT method");
        if (nshape < c.steps) fail(QDX_ERR_PARAM, "the shape function is shorter than the propagation");
        if (U->imag) fail(QDX_ERR_INCOMPAT, "OCT needs a real-time propagator");
        if (U->backward) fail(QDX_ERR_STATE, "the propagator must start in the forward direction");
        if (opts->operator_target ? !targets : !psi_t) fail(QDX_ERR_PARAM, "targets missing");
        for (int t = 0; t < c.nt; ++t) {
            if (!psi_i[t] || !psi_i[t]->ctx) fail(QDX_ERR_PARAM, "Missing initial wavefunction");
            if (!same_shape(psi_i[t], psi_i[0])) fail(QDX_ERR_INCOMPAT, "initial wave functions differ in shape");
            c.psi_i.push_back(psi_i[t]);
            if (opts->operator_target) {
                if (!targets[t]) fail(QDX_ERR_PARAM, "Missing target operator");
                if (!same_shape(targets[t]->buf, psi_i[0])) fail(QDX_ERR_INCOMPAT, "target operator shape");
                c.targets.push_back(targets[t]);
            } else {
                if (!psi_t[t] || !same_shape(psi_t[t], psi_i[0]))
                    fail(QDX_ERR_PARAM, "Missing target wavefunction");
                c.psi_t.push_back(psi_t[t]);
            }
        }
        if (!(U->grid == psi_i[0]->grid) || U->S != psi_i[0]->S || U->ctx != psi_i[0]->ctx)
            fail(QDX_ERR_INCOMPAT, "wave function does not match the propagator");
        const std::vector<qdx_ham *> hs = field_holders(U);
        if (hs.empty()) fail(QDX_ERR_PARAM, "Propagator has no hamiltonian");
        for (qdx_ham *h : hs)
            for (Term *t : dipole_terms(h)) c.dip.push_back(t);
        if (c.dip.empty()) fail(QDX_ERR_PARAM, "Given type of coupling operator not found in hamiltonian");
        if (laser_min_len(c.dip) < c.steps) fail(QDX_ERR_PARAM, "Clock values doesn't fit the laser field");
        c.lscale = c.dip.front()->scale;
        c.laser_dt = U->dt;
        c.Ef.assign(c.dip.front()->laser.begin(), c.dip.front()->laser.end());
        c.Ef.resize(laser_min_len(c.dip));
        c.Eb = c.Ef;   // _laserb = copy of _laserf (ProgOCT.cpp:939-942)
        c.push_field();
        c.coup.reset(coupling_operator(U, hs));
        if (gobbler && (!(gobbler->grid == U->grid) || gobbler->S != U->S || gobbler->ctx != U->ctx))
            fail(QDX_ERR_INCOMPAT, "gobbler does not match the propagator");

        // the backward-propagation buffer: nsteps + 1 functions per target
        size_t freeb = 0, total = 0;
        HIPCHK(hipMemGetInfo(&freeb, &total));
        const double need = double(c.nt) * double(c.steps + 1 + 3) * double(psi_i[0]->n()) * psi_i[0]->S * 16.0;
        if (need > 0.9 * double(freeb))
            fail(QDX_ERR_MEMORY, "OCT buffer (" + std::to_string(need / 1e9) + " GB) exceeds free device memory");
        c.membuf.resize(c.nt);
        for (int t = 0; t < c.nt; ++t)
            for (long s = 0; s <= c.steps; ++s) c.membuf[t].push_back(c.own.make(psi_i[0]));
        c.opwf = c.own.make(psi_i[0]);
        if (gobbler) c.gbuf = c.own.make(psi_i[0]);
        std::vector<qdx_wf *> phii(c.nt), phit(c.nt);
        for (int t = 0; t < c.nt; ++t) {
            phii[t] = c.own.make(psi_i[0]);
            phit[t] = c.own.make(psi_i[0]);
        }
        const int rowlen = QDX_OCT_ROW_FIXED + QDX_OCT_ROW_PER_TARGET * c.nt;
        int rows = 0;
        auto row = [&](int it) { return report ? report + (size_t)it * rowlen : nullptr; };

        // ProgOCT::Run (ProgOCT.cpp:1009-1092): the guess, then the iterations
        c.copy_set(phii, c.psi_i);
        c.laserobj = 0;
        c.forward(phii);
        c.report(phii, c.psi_t.empty() ? phii : c.psi_t, 0, row(0));
        rows = 1;
        double dT[3] = {1, 1, 1};
        double delta = 1;
        int it = 1;
        while (it <= opts->iterations && delta > opts->conv) {
            if ((opts->method == QDX_OCT_RABITZFB && it == 1) || opts->method != QDX_OCT_RABITZFB) {
                c.copy_set(phii, c.psi_i);
                c.copy_set(phit, opts->operator_target ? c.psi_i : c.psi_t);
            }
            c.laserobj = 0;
            if (opts->method == QDX_OCT_RABITZFB) c.iterate_fb(phii, phit, it);
            else c.iterate(phii, phit);
            c.laserobj *= U->dt;
            dT[0] = dT[1];
            dT[1] = dT[2];
            dT[2] = std::abs(c.report(phii, phit, it, row(it)));
            rows = it + 1;
            delta = std::abs(dT[0] - dT[1]) + std::abs(dT[1] - dT[2]);
            c.swap_fields();   // new <-> old
            ++it;
        }
        // the optimised field (QDng writes _laserb as the final laser)
        c.Ef = c.Eb;
        c.push_field();
        chk(qdx_ctx_synchronize(U->ctx));
        if (rows_done) *rows_done = rows;
        if (converged) *converged = delta <= opts->conv ? 1 : 0;
    });
}

}  // extern "C"
