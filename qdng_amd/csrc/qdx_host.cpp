// Host runtime of libqdx.so: handles, QDng's operator rules (bounds,
// multistate mirroring, Chebyshev order and coefficients, GSPO operator
// sequence) and the step drivers that enqueue the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "qdx_host_internal.hpp"


namespace {

qdx_wf *check_wf(qdx_wf *w) {
    if (!w || !w->ctx) fail(QDX_ERR_PARAM, "Invalid WaveFunction");
    return w;
}

// points per thread of the Chebyshev line kernels on this grid (see
// qdx_ctx_set_line_points).  reduced: the FP32 and compact Chebyshev
// kernels, which take 4 wherever the lines fit (3D 128^3: FP32 2.46 -> 2.12,
// compact 3.41 -> 2.91 ms per step; the FP64 kernels time the same at 4 and
// 8 on that grid)
int cheby_line_points(const qdx_ctx *ctx, const GridDev &g, int S, bool reduced = false) {
    if (g.nd < 2 || ctx->line_ppt == 8) return 8;
    if (S * (g.N[0] / 4) > 512) return 8;
    for (int d = 1; d < g.nd; ++d)
        if (g.N[d] > 2048) return 8;
    if (ctx->line_ppt == 4 || reduced) return 4;
    const long waves8 = (long)S * g.size / (8 * 64);
    return waves8 < 2 * 1024 ? 4 : 8;   // 256 CUs x 4 SIMDs
}

template <class R>
void cheby_step(qdx_prop *p, qdx_wf *psi, long step) {
    qdx_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const bool fp32 = std::is_same<R, float>::value;
    const GridDev g = p->grid.dev();
    qdx_ham *H = p->ham;
    HamDev<R> h;
    if constexpr (std::is_same<R, float>::value) h = H->dev.h32;
    else h = H->dev.h64;
    H->set_fields(h, step);
    LinePlan plans[kMaxDims];
    const int ppt = cheby_line_points(ctx, g, p->S, fp32 || p->cmp);
    for (int d = 0; d < p->grid.nd; ++d) plans[d] = ctx->plan_ppt(p->grid.N[d], fp32, ppt);
    R *A = p->phiA.as<R>(), *B = p->phiB.as<R>();
    R *X = p->grid.nd > 1 ? p->X.as<R>() : nullptr;
    ChebyCtl *ctl = p->ctl.as<ChebyCtl>();
    // partial-sum slots: [0],[1] alternate between orders, [2] holds the
    // sum |psi|^2 of the step start
    // (imaginary time compares two previous orders: three slots rotate)
    double *part = p->partials.as<double>();
    const int nslot = p->imag ? 3 : 2;
    double *slot[3] = {part, part + kMaxPartials, part + 3 * kMaxPartials};
    double *init = part + 2 * kMaxPartials;
    if (p->line_step) {
        KTimer t(ctx, QDX_K_CHEBY_ROW);
        // the whole-step kernel of a single-state line: 4 points per thread
        // up to 1024 points (a short line spreads over more lanes: 3.7-5.9
        // us per order against 5.2-7.4 at 8 for 128-1024 points), 8 above;
        // qdx_ctx_set_line_points 4 or 8 forces the choice (4 up to 1024)
        const int N0 = p->grid.N[0];
        const int lp = (ctx->line_ppt == 8 || N0 > 1024) ? 8 : 4;
        const LinePlan lplan = lp == 8 ? plans[0] : ctx->plan_ppt(N0, fp32, 4);
        Launch<R>::cheby_line(s, g, h, lplan, p->coefd.as<double>(), p->k_launch, p->cp.offset, p->cp.offset_i,
                              1.0 / p->cp.scaling, p->cp.prec, psi->ptr(), ctl, p->partials.as<double>());
        p->last_nb = 1;
        p->nb_init = 1;
        return;
    }
    const int S = p->S;
    int nb_init;
    {
        KTimer t(ctx, QDX_K_CHEBY_INIT);
        nb_init = Launch<R>::cheby_init(s, g, S, plans[0], psi->ptr(), A, nullptr, ctl, init, p->cmp);
    }

    ConvCheck cc;
    cc.norm0 = &ctl->norm0;
    // imaginary time: the relative test, marked by prec < 0, with norm0
    // pointing at the partials of the order before (set per order below)
    cc.prec = p->imag ? -p->cp.prec : p->cp.prec;
    // T = sum_d T_d: the kinetic parts along dimensions >= 1 of phi_{k-1} go
    // to X; the row kernel adds the dimension-0 part
    auto cols = [&](const R *pm1, const ConvCheck &c) {
        for (int d = 1; d < p->grid.nd; ++d) {
            KTimer t(ctx, QDX_K_LINE_PASS);
            Launch<R>::kin_pass(s, g, S, d, d > 1, plans[d], h, pm1, X, ctl, c, p->cmp);
        }
    };
    // exp(exponent * offset), exponent -i dt (real time) or -dt (imaginary);
    // the offset is complex with a negative imaginary potential
    const std::complex<double> ph =
        p->imag ? std::complex<double>(std::exp(-p->dt * p->cp.offset), 0.0)
                : std::exp(std::complex<double>(0, -1) * p->dt * std::complex<double>(p->cp.offset, p->cp.offset_i));
    ChebyOrder ord;
    ord.offset = p->cp.offset;
    ord.offset_i = p->cp.offset_i;
    ord.inv_scaling = 1.0 / p->cp.scaling;
    ord.prec = p->cp.prec;
    const std::complex<double> c0 = p->cp.coeff[0] * ph, c1 = p->cp.coeff[1] * ph;
    ord.k = 1;
    ord.c0r = c0.real();
    ord.c0i = c0.imag();
    ord.ckr = c1.real();
    ord.cki = c1.imag();
    cols(A, cc);   // k_prev = 0: no test
    {
        KTimer t(ctx, QDX_K_CHEBY_ROW);
        Launch<R>::cheby_row(s, g, h, plans[0], ord, cc, p->imag, A, B, X, psi->ptr(), ctl, slot[1 % nslot], init,
                             nb_init, p->cmp);
    }
    int nb_prev = 0;
    for (int k = 2; k < p->k_launch; ++k) {
        const std::complex<double> ck = p->cp.coeff[k] * ph;
        ord.k = k;
        ord.ckr = ck.real();
        ord.cki = ck.imag();
        // fixed-length series (no break test): psi is updated every third
        // order and at the last one with the terms since the previous update
        ord.nterm = 1;
        if (!(p->cp.prec > 0)) {
            const int m = (k - 1) % 3;
            ord.nterm = m == 0 ? 3 : (k == p->k_launch - 1 ? m : 0);
            const std::complex<double> c1 = p->cp.coeff[k - 1] * ph, c2 = p->cp.coeff[k - 2] * ph;
            ord.cm1r = c1.real();
            ord.cm1i = c1.imag();
            ord.cm2r = c2.real();
            ord.cm2i = c2.imag();
        }
        cc.k_prev = k - 1;
        cc.part = slot[(k - 1) % nslot];
        if (p->imag) cc.norm0 = slot[(k - 2) % nslot];
        cc.nb = nb_prev;
        R *pm1 = (k - 1) % 2 == 0 ? A : B;
        R *pm2 = k % 2 == 0 ? A : B;
        cols(pm1, cc);
        KTimer t(ctx, QDX_K_CHEBY_ROW);
        nb_prev = Launch<R>::cheby_row(s, g, h, plans[0], ord, cc, p->imag, pm1, pm2, X, psi->ptr(), ctl,
                                       slot[k % nslot], init, nb_init, p->cmp);
    }
    p->last_nb = nb_prev;
    p->nb_init = nb_init;
}

// --- GSPO -------------------------------------------------------------------
struct SplitOp {
    bool kinetic = false;
    qdx_ham *op = nullptr;
    double tau = 0;
};

std::vector<SplitOp> gspo_sequence(qdx_prop *p) {
    // OGSPO.cpp:119-138: ops n-1..1 with dt/2, then 0 with dt, then 1..n-1 with dt/2
    std::vector<SplitOp> seq;
    const int n = (int)p->ops.size();
    auto mk = [&](int i, double sf) {
        SplitOp o;
        o.op = p->ops[i];
        o.kinetic = o.op->terms[0].kind == T_KIN;
        o.tau = p->dt * sf;
        return o;
    };
    for (int i = n - 1; i > 0; --i) seq.push_back(mk(i, 0.5));
    for (int i = 0; i < n; ++i) seq.push_back(mk(i, i == 0 ? 1.0 : 0.5));
    return seq;
}

// exp(-i tau T_d(k)) for every state and dimension of a kinetic split
// operator (layout of k_col's ek: state-major, dimensions in order, (re, im)
// pairs), from the operator's own device tables T_d; a state without a
// kinetic term gets 1.  Built once per (operator, tau).
const double *kinetic_exp_table(qdx_prop *p, qdx_ham *op, double tau) {
    const auto key = std::make_pair((const qdx_ham *)op, tau);
    auto it = p->kinexp.find(key);
    if (it != p->kinexp.end()) return it->second->as<double>();
    const Grid &g = p->grid;
    long tot = 0;
    for (int d = 0; d < g.nd; ++d) tot += g.N[d];
    std::vector<double> e(2 * (size_t)op->S * tot);
    const HamDev<double> &h = op->dev.h64;
    std::vector<double> t;
    for (int s = 0; s < op->S; ++s) {
        long off = 0;
        for (int d = 0; d < g.nd; ++d) {
            t.assign(g.N[d], 0.0);
            if (h.tk[s][0]) HIPCHK(hipMemcpy(t.data(), h.tk[s][d], g.N[d] * sizeof(double), hipMemcpyDeviceToHost));
            for (int k = 0; k < g.N[d]; ++k) {
                e[2 * (s * tot + off + k)] = std::cos(tau * t[k]);
                e[2 * (s * tot + off + k) + 1] = -std::sin(tau * t[k]);
            }
            off += g.N[d];
        }
    }
    auto buf = std::make_unique<DevBuf>(e.size() * sizeof(double));
    HIPCHK(hipMemcpy(buf->p, e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice));
    return p->kinexp.emplace(key, std::move(buf)).first->second->as<double>();
}

template <class R>
XSplitDev<R> xsplit(qdx_prop *p, qdx_ham *op, double tau, long step) {
    XSplitDev<R> x;
    x.S = op->S;
    x.tau = tau;
    x.coupled = 0;
    HamDev<R> h;
    if constexpr (std::is_same<R, float>::value) h = op->dev.h32;
    else h = op->dev.h64;
    op->set_fields(h, step);
    bool dipole = false;
    for (int i = 0; i < op->S; ++i)
        for (int j = 0; j < op->S; ++j) {
            x.V[i][j] = h.W[i][j];
            x.mu[i][j] = h.mu[i][j];
            x.Edip[i][j] = h.Edip[i][j];
            if (h.mu[i][j]) dipole = true;
            if (i != j && (h.W[i][j] || h.mu[i][j])) x.coupled = 1;
        }
    // a real-time operator without a field is the same unitary at every
    // step: its per-point table is computed once (the operator's terms are
    // fixed once it is in use)
    if (!p->imag && !dipole && (!x.coupled || op->S == 2)) {
        const auto key = std::make_tuple((const qdx_ham *)op, tau, std::is_same<R, float>::value);
        auto it = p->unitaries.find(key);
        if (it == p->unitaries.end()) {
            const GridDev g = p->grid.dev();
            const size_t ncomp = x.coupled ? 3 : (size_t)op->S;
            auto buf = std::make_unique<DevBuf>(ncomp * (size_t)g.size * 2 * sizeof(double));
            Launch<R>::gspo_unitary(p->ctx->stream, g, op->S, x, buf->as<double>());
            it = p->unitaries.emplace(key, std::move(buf)).first;
        }
        x.U = it->second->template as<double>();
        x.n = p->grid.size();
    }
    return x;
}

// Strided passes of the full kinetic factor on grids with nd >= 2 (mode 2:
// T(k)/N, mode 3: exp(-i tau T(k))/N) on data transformed along dimension 0:
// forward transforms along dimensions 1 .. nd-2, the last dimension's
// transform, factor and inverse, then the inverse transforms back down to
// dimension 1 (QDng's N-dimensional transform, OGSPO.cpp:114-139).
template <class F>
void full_t_passes(int nd, int mode, F col) {
    for (int d = 1; d < nd - 1; ++d) col(d, 0);
    col(nd - 1, mode);
    for (int d = nd - 2; d >= 1; --d) col(d, 1);
}

template <class R>
void gspo_step(qdx_prop *p, R *psi, long step) {
    qdx_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const bool fp32 = std::is_same<R, float>::value;
    const GridDev g = p->grid.dev();
    LinePlan plans[kMaxDims];
    // 4 points per thread where the grid is too small for two waves per
    // SIMD at 8 (the rule of the Chebyshev line kernels)
    const int ppt = cheby_line_points(ctx, g, p->S);
    for (int d = 0; d < p->grid.nd; ++d) plans[d] = ctx->plan_ppt(p->grid.N[d], fp32, ppt);
    auto seq = gspo_sequence(p);
    XGroupDev<R> pending;
    pending.n = 0;
    size_t i = 0;
    while (i < seq.size()) {
        const SplitOp &o = seq[i];
        if (!o.kinetic) {
            if (pending.n == kMaxXOps) {
                Launch<R>::gspo_xonly(s, g, p->S, pending, psi, p->imag);
                pending.n = 0;
            }
            pending.op[pending.n++] = xsplit<R>(p, o.op, o.tau, step);
            ++i;
            continue;
        }
        // kinetic split operator: collect the x-space operators that follow it
        XGroupDev<R> post;
        post.n = 0;
        size_t j = i + 1;
        while (j < seq.size() && !seq[j].kinetic && post.n < kMaxXOps) {
            post.op[post.n++] = xsplit<R>(p, seq[j].op, seq[j].tau, step);
            ++j;
        }
        HamDev<R> kin;
        if constexpr (std::is_same<R, float>::value) kin = o.op->dev.h32;
        else kin = o.op->dev.h64;
        if (p->grid.nd == 1) {
            KTimer t(ctx, QDX_K_GSPO_ROW);
            Launch<R>::gspo_1d(s, g, p->S, plans[0], kin, o.tau, pending, post, psi, p->imag);
        } else {
            R *X = p->X.as<R>();
            {
                KTimer t(ctx, QDX_K_GSPO_ROW);
                Launch<R>::gspo_row_pre(s, g, p->S, plans[0], pending, psi, X, p->imag);
            }
            const double *ek = p->imag ? nullptr : kinetic_exp_table(p, o.op, o.tau);
            auto col = [&](int dim, int mode) {
                KTimer t(ctx, QDX_K_LINE_PASS);
                Launch<R>::col_pass(s, g, p->S, dim, mode, plans[dim], kin, o.tau, X, nullptr, ConvCheck(), p->imag,
                                    ek);
            };
            full_t_passes(p->grid.nd, 3, col);
            {
                KTimer t(ctx, QDX_K_GSPO_ROW);
                Launch<R>::gspo_row_post(s, g, p->S, plans[0], post, X, psi, p->imag);
            }
        }
        pending.n = 0;
        i = j;
    }
    if (pending.n > 0) Launch<R>::gspo_xonly(s, g, p->S, pending, psi, p->imag);
}

// GSPO runs of several steps on grids with nd >= 2 whose sequence holds one
// kinetic split operator (x-space groups before and after it): psi stays
// transformed along dimension 0 in X between the steps, and one row pass
// ends step n and starts step n + 1 (post of n, pre of n + 1).  Per step: the
// strided passes and one row pass, plus one row pass at the start.
struct GspoFused {
    bool ok = false;
    size_t kpos = 0;   // index of the kinetic operator in the sequence
};

GspoFused gspo_fusable(qdx_prop *p) {
    GspoFused f;
    if (p->kind != qdx_prop::GSPO || p->grid.nd < 2) return f;
    const auto seq = gspo_sequence(p);
    int nkin = 0;
    for (size_t i = 0; i < seq.size(); ++i)
        if (seq[i].kinetic) {
            ++nkin;
            f.kpos = i;
        }
    f.ok = nkin == 1 && f.kpos <= (size_t)kMaxXOps && seq.size() - f.kpos - 1 <= (size_t)kMaxXOps;
    return f;
}

// steps first .. first + nsteps - 1 of a fusable GSPO run on psi (complex R);
// obs_part/obs: when given, the Reporter sums of the state before steps
// 1 .. nsteps - 1 go to obs + i * per_step (as qdx_prop_run lays them out)
template <class R>
void gspo_run_fused(qdx_prop *p, R *psi, const R *psi0, long first, long nsteps, double *obs, size_t per_step) {
    qdx_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const bool fp32 = std::is_same<R, float>::value;
    const GridDev g = p->grid.dev();
    LinePlan plans[kMaxDims];
    const int ppt = cheby_line_points(ctx, g, p->S);
    for (int d = 0; d < p->grid.nd; ++d) plans[d] = ctx->plan_ppt(p->grid.N[d], fp32, ppt);
    const auto seq = gspo_sequence(p);
    const size_t kpos = gspo_fusable(p).kpos;
    auto group = [&](size_t a, size_t b, long step) {
        XGroupDev<R> grp;
        grp.n = 0;
        for (size_t i = a; i < b; ++i) grp.op[grp.n++] = xsplit<R>(p, seq[i].op, seq[i].tau, step);
        return grp;
    };
    HamDev<R> kin;
    if constexpr (std::is_same<R, float>::value) kin = seq[kpos].op->dev.h32;
    else kin = seq[kpos].op->dev.h64;
    const double tau = seq[kpos].tau;
    const double *ek = p->imag ? nullptr : kinetic_exp_table(p, seq[kpos].op, tau);
    R *X = p->X.as<R>();
    double *opart = obs ? p->partials.as<double>() : nullptr;
    {
        KTimer t(ctx, QDX_K_GSPO_ROW);
        Launch<R>::gspo_row_pre(s, g, p->S, plans[0], group(0, kpos, first), psi, X, p->imag);
    }
    for (long i = 0; i < nsteps; ++i) {
        const long step = first + i;
        auto col = [&](int dim, int mode) {
            KTimer t(ctx, QDX_K_LINE_PASS);
            Launch<R>::col_pass(s, g, p->S, dim, mode, plans[dim], kin, tau, X, nullptr, ConvCheck(), p->imag, ek);
        };
        full_t_passes(p->grid.nd, 3, col);
        const XGroupDev<R> post = group(kpos + 1, seq.size(), step);
        if (i == nsteps - 1) {
            KTimer t(ctx, QDX_K_GSPO_ROW);
            Launch<R>::gspo_row_post(s, g, p->S, plans[0], post, X, psi, p->imag);
            break;
        }
        int nb;
        {
            KTimer t(ctx, QDX_K_GSPO_ROW);
            nb = Launch<R>::gspo_row_bridge(s, g, p->S, plans[0], post, group(0, kpos, step + 1), X, psi0, opart,
                                            p->imag);
        }
        if (opart) {
            double *o = obs + (i + 1) * per_step;
            launch_sum_partials(s, opart, nb, p->S, o);                         // <psi|psi>
            launch_sum_partials(s, opart + 2 * p->S * nb, nb, p->S, o + 2 * p->S);   // <psi|psi0>
        }
    }
}

// |norm0 - accnorm| / norm0 of the last executed order of the last step
// (imaginary time: |accnorm - accnorm of the order before| / the latter),
// from the partial sums the kernels stored
double cheby_last_nconv(qdx_prop *p) {
    ChebyCtl c;
    HIPCHK(hipMemcpy(&c, p->ctl.p, sizeof(c), hipMemcpyDeviceToHost));
    if (c.last_order < (p->imag ? 3 : 2)) return 0.0;
    const double *part = p->partials.as<double>();
    // slot of order k, as cheby_step lays them out
    auto slot = [&](int k) {
        const int i = k % (p->imag ? 3 : 2);
        return part + (i == 2 ? 3 : i) * kMaxPartials;
    };
    auto sum = [&](const double *src, int n) {
        std::vector<double> a(n);
        HIPCHK(hipMemcpy(a.data(), src, a.size() * sizeof(double), hipMemcpyDeviceToHost));
        double t = 0;
        for (double v : a) t += v;
        return t;
    };
    const double acc = sum(slot(c.last_order), p->last_nb);
    const double ref = p->imag ? sum(slot(c.last_order - 1), p->last_nb) : sum(part + 2 * kMaxPartials, p->nb_init);
    return std::fabs(ref - acc) / ref;
}

void sil_step(qdx_prop *p, qdx_wf *psi, long step);
void arnoldi_step(qdx_prop *p, qdx_wf *psi, long step);

void prop_step_forward(qdx_prop *p, qdx_wf *psi, long step);

// OPropagator::Backward() in real time (exponent +i dt, with the operators
// conjugated, OPropagator.h:82-100): exp(+i H* dt) psi = conj(exp(-i H dt)
// conj(psi)), the forward step between two conjugations (exact: conjugation
// commutes with every rounding of the forward step)
void prop_step(qdx_prop *p, qdx_wf *psi, long step) {
    if (!p->backward) return prop_step_forward(p, psi, step);
    const bool g32 = p->kind == qdx_prop::GSPO && p->precision == QDX_FP32_BULK;
    void *d = g32 ? static_cast<void *>(p->psi32.as<float>()) : static_cast<void *>(psi->ptr());
    const long n = (long)p->S * psi->n();
    launch_conj(p->ctx->stream, d, n, g32);
    prop_step_forward(p, psi, step);
    launch_conj(p->ctx->stream, d, n, g32);
}

void prop_step_forward(qdx_prop *p, qdx_wf *psi, long step) {
    if (p->kind == qdx_prop::SIL) {
        sil_step(p, psi, step);
        return;
    }
    if (p->kind == qdx_prop::ARNOLDI) {
        arnoldi_step(p, psi, step);
        return;
    }
    if (p->kind == qdx_prop::CHEBY) {
        // QDX_COMPACT: FP64 arithmetic (compact terms when p->cmp)
        if (p->precision == QDX_FP32_BULK) cheby_step<float>(p, psi, step);
        else cheby_step<double>(p, psi, step);
        // record executed order count of this step
    } else {
        if (p->precision == QDX_FP32_BULK) gspo_step<float>(p, p->psi32.as<float>(), step);
        else gspo_step<double>(p, psi->ptr(), step);
    }
}

void check_prop_wf(qdx_prop *p, qdx_wf *psi) {
    check_wf(psi);
    if (psi->ctx != p->ctx) fail(QDX_ERR_INCOMPAT, "wave function belongs to another context");
    if (!(psi->grid == p->grid) || psi->S != p->S)
        fail(QDX_ERR_INCOMPAT, "GridPotential parameters are not compatible with wave function");
}

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

const char *qdx_last_error(void) { return g_err.c_str(); }
const char *qdx_version(void) { return "qdx 0.1 (gfx950)"; }

int qdx_ctx_create(int device, qdx_ctx **out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "out is NULL");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) fail(QDX_ERR_DEVICE, "no HIP device available");
        if (device < 0 || device >= n) fail(QDX_ERR_PARAM, "device index out of range");
        HIPCHK(hipSetDevice(device));
        auto c = std::make_unique<qdx_ctx>();
        c->device = device;
        // device scratch of qdx_ham_expec_noalloc, looked up once per device
        qdx::noalloc_scratch_setup(device);
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->partials.alloc(kMaxPartials * sizeof(double));
        c->scal.alloc(4096 * sizeof(double));
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&c->pinned), 4096 * sizeof(double), hipHostMallocDefault));
        *out = c.release();
    });
}

int qdx_ctx_profile(qdx_ctx *ctx, int32_t on) {
    return guarded([&] {
        if (!ctx) fail(QDX_ERR_PARAM, "ctx is NULL");
        if (ctx->profiling) ctx->fold();
        ctx->profiling = on != 0;
        if (ctx->profiling)
            for (int c = 0; c < QDX_NUM_KERNEL_CLASSES; ++c) {
                ctx->acc_ms[c] = 0;
                ctx->acc_n[c] = 0;
            }
    });
}

int qdx_ctx_kernel_time(qdx_ctx *ctx, int32_t kernel_class, int64_t *launches, double *total_ms) {
    return guarded([&] {
        if (!ctx || !launches || !total_ms) fail(QDX_ERR_PARAM, "NULL argument");
        if (kernel_class < 0 || kernel_class >= QDX_NUM_KERNEL_CLASSES) fail(QDX_ERR_PARAM, "unknown kernel class");
        ctx->fold();
        *launches = ctx->acc_n[kernel_class];
        *total_ms = ctx->acc_ms[kernel_class];
    });
}

int qdx_ctx_set_radix_plan(qdx_ctx *ctx, int32_t precision, int32_t L, const int32_t *radix, int32_t nst) {
    return guarded([&] {
        if (!ctx || !radix) fail(QDX_ERR_PARAM, "NULL argument");
        if (precision != QDX_FP64 && precision != QDX_FP32_BULK) fail(QDX_ERR_PARAM, "unknown precision");
        if (nst < 1 || nst > kMaxStages) fail(QDX_ERR_PARAM, "1 to 16 radix stages");
        radix_plan(L);   // a valid line length
        std::vector<int> r(radix, radix + nst);
        if (!radix_ok(L, r)) fail(QDX_ERR_PARAM, "radix plan does not fit the line length");
        ctx->radix[{L, precision == QDX_FP32_BULK}] = r;
    });
}

int qdx_ctx_get_radix_plan(qdx_ctx *ctx, int32_t precision, int32_t L, int32_t *radix, int32_t *nst) {
    return guarded([&] {
        if (!ctx || !radix || !nst) fail(QDX_ERR_PARAM, "NULL argument");
        if (precision != QDX_FP64 && precision != QDX_FP32_BULK && precision != QDX_COMPACT)
            fail(QDX_ERR_PARAM, "unknown precision");
        const std::vector<int> r = ctx->radix_of(L, precision == QDX_FP32_BULK);
        for (size_t i = 0; i < r.size(); ++i) radix[i] = r[i];
        *nst = (int32_t)r.size();
    });
}

int qdx_ctx_set_line_points(qdx_ctx *ctx, int32_t points) {
    return guarded([&] {
        if (!ctx) fail(QDX_ERR_PARAM, "ctx is NULL");
        if (points != 0 && points != 4 && points != 8) fail(QDX_ERR_PARAM, "points per thread: 0, 4 or 8");
        ctx->line_ppt = points;
    });
}

int qdx_ctx_get_line_points(qdx_ctx *ctx, int32_t *points) {
    return guarded([&] {
        if (!ctx || !points) fail(QDX_ERR_PARAM, "NULL argument");
        *points = ctx->line_ppt;
    });
}

int qdx_ctx_synchronize(qdx_ctx *ctx) {
    return guarded([&] {
        if (!ctx) fail(QDX_ERR_PARAM, "ctx is NULL");
        HIPCHK(hipStreamSynchronize(ctx->stream));
    });
}

int qdx_ctx_destroy(qdx_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        // children keep the context alive until the last one is destroyed
        ctx->closing = true;
        if (ctx->refs == 0) ctx_free(ctx);
    });
}

int qdx_wf_create(qdx_ctx *ctx, const qdx_grid *grid, int32_t nstates, qdx_wf **out) {
    return guarded([&] {
        if (!ctx || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (nstates < 1 || nstates > kMaxStates) fail(QDX_ERR_PARAM, "number of states must be 1..4");
        auto w = std::make_unique<qdx_wf>();
        w->ctx = ctx;
        w->grid = make_grid(grid);
        w->S = nstates;
        w->d.alloc((size_t)nstates * w->grid.size() * 2 * sizeof(double));
        HIPCHK(hipMemsetAsync(w->d.p, 0, w->d.bytes, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ctx_acquire(ctx);
        *out = w.release();
    });
}

int qdx_wf_upload(qdx_wf *wf, int32_t state, const double *c128) {
    return guarded([&] {
        check_wf(wf);
        if (!c128) fail(QDX_ERR_PARAM, "NULL array");
        const size_t one = wf->n() * 2 * sizeof(double);
        if (state >= wf->S) fail(QDX_ERR_PARAM, "state index out of range");
        double *dst = state < 0 ? wf->ptr() : wf->ptr() + (size_t)state * wf->n() * 2;
        HIPCHK(hipMemcpyAsync(dst, c128, state < 0 ? one * wf->S : one, hipMemcpyHostToDevice, wf->ctx->stream));
        HIPCHK(hipStreamSynchronize(wf->ctx->stream));
    });
}

int qdx_wf_download(qdx_wf *wf, int32_t state, double *c128) {
    return guarded([&] {
        check_wf(wf);
        if (!c128) fail(QDX_ERR_PARAM, "NULL array");
        const size_t one = wf->n() * 2 * sizeof(double);
        if (state >= wf->S) fail(QDX_ERR_PARAM, "state index out of range");
        const double *src = state < 0 ? wf->ptr() : wf->ptr() + (size_t)state * wf->n() * 2;
        HIPCHK(hipMemcpyAsync(c128, src, state < 0 ? one * wf->S : one, hipMemcpyDeviceToHost, wf->ctx->stream));
        HIPCHK(hipStreamSynchronize(wf->ctx->stream));
    });
}

int qdx_wf_copy(qdx_wf *dst, const qdx_wf *src) {
    return guarded([&] {
        check_wf(dst);
        if (!src || !(src->grid == dst->grid) || src->S != dst->S)
            fail(QDX_ERR_INCOMPAT, "wave functions differ in grid or states");
        HIPCHK(hipMemcpyAsync(dst->d.p, src->d.p, dst->d.bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
        HIPCHK(hipStreamSynchronize(dst->ctx->stream));
    });
}

static void overlap_states(qdx_wf *a, qdx_wf *b, std::vector<double> &per) {
    qdx_ctx *c = a->ctx;
    launch_overlap(c->stream, a->ptr(), b->ptr(), a->n(), a->S, c->partials.as<double>(), c->scal.as<double>());
    HIPCHK(hipMemcpyAsync(c->pinned, c->scal.p, 2 * a->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    per.assign(c->pinned, c->pinned + 2 * a->S);
}

int qdx_wf_norm(qdx_wf *wf, double *norm, double *per_state_pop) {
    return guarded([&] {
        check_wf(wf);
        std::vector<double> per;
        overlap_states(wf, wf, per);
        const double w = wf->grid.weight();
        double tot = 0;
        for (int s = 0; s < wf->S; ++s) {   // WFMultistate.cpp:158-174
            const double p = per[2 * s] * w;
            if (per_state_pop) per_state_pop[s] = p;
            tot += p;
        }
        if (norm) *norm = std::sqrt(tot);
        if (!std::isfinite(tot)) fail(QDX_ERR_OVERFLOW, "Norm is not a number");
    });
}

int qdx_wf_overlap(qdx_wf *a, qdx_wf *b, double out[2]) {
    return guarded([&] {
        check_wf(a);
        check_wf(b);
        if (!(a->grid == b->grid) || a->S != b->S) fail(QDX_ERR_INCOMPAT, "wave functions differ in grid or states");
        std::vector<double> per;
        overlap_states(a, b, per);
        const double w = a->grid.weight();
        double re = 0, im = 0;
        for (int s = 0; s < a->S; ++s) {
            re += per[2 * s] * w;
            im += per[2 * s + 1] * w;
        }
        out[0] = re;
        out[1] = im;
    });
}

int qdx_wf_normalize(qdx_wf *wf) {
    return guarded([&] {
        check_wf(wf);
        std::vector<double> per;
        overlap_states(wf, wf, per);
        double tot = 0;
        for (int s = 0; s < wf->S; ++s) tot += per[2 * s] * wf->grid.weight();
        const double f = 1.0 / std::sqrt(tot);
        launch_scale_axpy(wf->ctx->stream, wf->ptr(), nullptr, (long)wf->S * wf->n(), f, 0.0, 0.0);
    });
}

// per-state results of launch_wnorm (in ctx->scal) to the host
static void wnorm_fetch(qdx_wf *wf, double factor, double *per_state) {
    qdx_ctx *c = wf->ctx;
    HIPCHK(hipMemcpyAsync(c->pinned, c->scal.p, 2 * wf->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int s = 0; s < wf->S; ++s) per_state[s] = c->pinned[2 * s] * factor;
}

int qdx_wf_expec_position(qdx_wf *wf, int32_t dim, double *per_state) {
    return guarded([&] {
        check_wf(wf);
        if (!per_state) fail(QDX_ERR_PARAM, "NULL array");
        if (dim < 0 || dim >= wf->grid.nd) fail(QDX_ERR_PARAM, "dimension out of range");
        qdx_ctx *c = wf->ctx;
        launch_wnorm(c->stream, wf->ptr(), wf->n(), wf->S, wf->grid.dev(), dim, wf->grid.xmin[dim], wf->grid.dx(dim),
                     nullptr, c->partials.as<double>(), c->scal.as<double>());
        wnorm_fetch(wf, wf->grid.weight(), per_state);
    });
}

int qdx_wf_expec_momentum(qdx_wf *wf, int32_t dim, double *per_state) {
    return guarded([&] {
        check_wf(wf);
        if (!per_state) fail(QDX_ERR_PARAM, "NULL array");
        if (dim < 0 || dim >= wf->grid.nd) fail(QDX_ERR_PARAM, "dimension out of range");
        qdx_ctx *c = wf->ctx;
        const size_t bytes = (size_t)wf->S * wf->n() * 2 * sizeof(double);
        if (c->wfx.bytes < bytes) c->wfx.alloc(bytes);
        const std::vector<double> k = kspace_d1(wf->grid, dim);
        if (c->ktab.bytes < k.size() * sizeof(double)) c->ktab.alloc(k.size() * sizeof(double));
        HIPCHK(hipMemcpyAsync(c->ktab.p, k.data(), k.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        const GridDev g = wf->grid.dev();
        Launch<double>::line_fft(c->stream, g, wf->S, dim, c->plan(wf->grid.N[dim], false), wf->ptr(),
                                 c->wfx.as<double>());
        launch_wnorm(c->stream, c->wfx.as<double>(), wf->n(), wf->S, g, dim, 0.0, 0.0, c->ktab.as<double>(),
                     c->partials.as<double>(), c->scal.as<double>());
        // (the host vector k stays alive until the copy has run: synchronised here)
        wnorm_fetch(wf, wf->grid.weight() / wf->grid.N[dim], per_state);
    });
}

int qdx_wf_axpy(qdx_wf *y, double re, double im, const qdx_wf *x) {
    return guarded([&] {
        check_wf(y);
        check_wf(const_cast<qdx_wf *>(x));
        if (y->ctx != x->ctx) fail(QDX_ERR_INCOMPAT, "wave functions on different contexts");
        if (y->S != x->S || y->n() != x->n()) fail(QDX_ERR_INCOMPAT, "wave functions differ in size");
        launch_scale_axpy(y->ctx->stream, y->ptr(), const_cast<qdx_wf *>(x)->ptr(), (long)y->S * y->n(), 1.0, re, im);
    });
}

int qdx_wf_info(const qdx_wf *wf, qdx_grid *grid, int32_t *nstates) {
    return guarded([&] {
        if (!wf || !wf->ctx || !grid || !nstates) fail(QDX_ERR_PARAM, "NULL argument");
        *grid = qdx_grid{};
        grid->ndims = wf->grid.nd;
        for (int d = 0; d < wf->grid.nd; ++d) {
            grid->N[d] = wf->grid.N[d];
            grid->xmin[d] = wf->grid.xmin[d];
            grid->xmax[d] = wf->grid.xmax[d];
        }
        *nstates = wf->S;
    });
}

int qdx_wf_destroy(qdx_wf *wf) {
    return guarded([&] {
        if (!wf) return;
        qdx_ctx *c = wf->ctx;
        (void)hipStreamSynchronize(c->stream);
        delete wf;
        ctx_release(c);
    });
}

int qdx_ham_create(qdx_ctx *ctx, const qdx_grid *grid, int32_t nstates, int32_t multistate, int32_t unity,
                   qdx_ham **out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "out is NULL");
        if (nstates < 1 || nstates > kMaxStates) fail(QDX_ERR_PARAM, "number of states must be 1..4");
        if (!multistate && nstates != 1) fail(QDX_ERR_PARAM, "a Sum Hamiltonian acts on one state");
        auto h = std::make_unique<qdx_ham>();
        h->ctx = ctx;
        h->grid = make_grid(grid);
        h->S = nstates;
        h->multistate = multistate != 0;
        h->unity = unity != 0;
        ctx_acquire(ctx);
        *out = h.release();
    });
}

static void check_elem(qdx_ham *h, int row, int col) {
    if (!h) fail(QDX_ERR_PARAM, "Hamiltonian is NULL");
    h->check_open();
    if (row < 0 || col < 0 || row >= h->S || col >= h->S) fail(QDX_ERR_OVERFLOW, "Multistate index out of range");
}

int qdx_ham_add_kinetic(qdx_ham *h, int32_t state, const double *mass) {
    return guarded([&] {
        check_elem(h, state, state);
        if (!mass) fail(QDX_ERR_PARAM, "mass is NULL");
        Term t;
        t.kind = T_KIN;
        t.row = t.col = state;
        bool eff = false;
        for (int d = 0; d < h->grid.nd; ++d) {   // OGridNablaSq.cpp:46-55
            if (mass[d] == 0) fail(QDX_ERR_PARAM, "Zero mass defined");
            if (mass[d] > 0) eff = true;
            t.mass.push_back(mass[d]);
        }
        if (!eff) fail(QDX_ERR_PARAM, "Nabla^2 Operator is empty (no masses defined)");
        h->terms.push_back(t);
    });
}

int qdx_ham_add_potential(qdx_ham *h, int32_t row, int32_t col, const double *V) {
    return guarded([&] {
        check_elem(h, row, col);
        if (!V) fail(QDX_ERR_PARAM, "V is NULL");
        Term t;
        t.kind = T_POT;
        t.row = row;
        t.col = col;
        t.v.assign(V, V + h->grid.size());
        for (double x : t.v)
            if (!std::isfinite(x)) fail(QDX_ERR_PARAM, "potential holds non-finite values");
        h->terms.push_back(t);
    });
}

// OGobbler (OGobbler.cpp:32-204): Butterworth edges per dimension, the low
// pass 1 / (1 + (i/x)^(2 order)) and the high pass 1 / (1 + (x/i)^(2 order))
// over the grid index i with the edge x in grid points
// (math_functions.h:40-69).  Start value 1, both passes multiply (nip = 0,
// a real potential); start value 0, the high pass adds (nip = 1, the
// negative imaginary potential -i v).  Times gain at the end.
int qdx_ham_add_gobbler(qdx_ham *h, int32_t row, int32_t col, const double *lp, const double *rp, int32_t order,
                        double gain, int32_t nip) {
    return guarded([&] {
        check_elem(h, row, col);
        if (!lp || !rp) fail(QDX_ERR_PARAM, "lp / rp is NULL");
        if (order < 1) fail(QDX_ERR_PARAM, "Filter order to small");
        if (!std::isfinite(gain)) fail(QDX_ERR_PARAM, "gain is not finite");
        if (nip && (h->multistate || row != col))
            fail(QDX_ERR_INCOMPAT, "negative imaginary potential: single-state (Sum) Hamiltonians only");
        const Grid &g = h->grid;
        const long n = g.size();
        Term t;
        t.kind = nip ? T_NIP : T_POT;
        t.row = row;
        t.col = col;
        t.fixed = true;
        t.fmax = {gain, 0.0};
        t.fmin = nip ? std::complex<double>(0.0, -gain) : std::complex<double>(0.0, 0.0);
        t.v.assign(n, nip ? 0.0 : 1.0);
        long stride = 1;
        for (int d = 0; d < g.nd; ++d) {
            const int Nd = g.N[d];
            for (int pass = 0; pass < 2; ++pass) {
                const double e = pass == 0 ? lp[d] : rp[d];
                if (std::isnan(e)) continue;
                const double x = (e - g.xmin[d]) / g.dx(d);
                // int(x) > N or int(x) < 0 as in OGobbler::_Init, without the cast
                if (!std::isfinite(x) || x >= Nd + 1.0 || x <= -1.0)
                    fail(QDX_ERR_PARAM, "Center of filter edge out of the grid");
                std::vector<double> f(Nd);
                for (int i = 0; i < Nd; ++i) {
                    const double r = pass == 0 ? double(i) / x : x / double(i);
                    f[i] = 1 / (1 + std::pow(r, 2 * double(order)));
                }
                const bool add = pass == 1 && nip;
                for (long j = 0; j < n; ++j) {
                    const double fv = f[(j / stride) % Nd];
                    t.v[j] = add ? t.v[j] + fv : t.v[j] * fv;
                }
            }
            stride *= Nd;
        }
        if (gain != 1)
            for (double &v : t.v) v *= gain;
        for (double v : t.v)
            if (std::isnan(v)) fail(QDX_ERR_PARAM, "Gobbler filter holds NaN (an edge at the first grid point)");
        h->terms.push_back(t);
    });
}

int qdx_ham_add_dipole(qdx_ham *h, int32_t row, int32_t col, const double *mu, const double *laser, int32_t nt,
                       double scale) {
    return guarded([&] {
        check_elem(h, row, col);
        if (!mu || !laser) fail(QDX_ERR_PARAM, "NULL array");
        if (nt < 1) fail(QDX_ERR_PARAM, "Laser field has zero size");
        Term t;
        t.kind = T_DIP;
        t.row = row;
        t.col = col;
        t.v.assign(mu, mu + h->grid.size());
        t.laser.assign(laser, laser + nt);
        t.scale = scale;
        h->terms.push_back(t);
    });
}

int qdx_ham_bounds(qdx_ham *h, double *emin, double *emax) {
    return guarded([&] {
        if (!h || !emin || !emax) fail(QDX_ERR_PARAM, "NULL argument");
        h->bounds(*emin, *emax);
    });
}

int qdx_ham_bounds_c(qdx_ham *h, double emin[2], double emax[2]) {
    return guarded([&] {
        if (!h || !emin || !emax) fail(QDX_ERR_PARAM, "NULL argument");
        h->bounds(emin[0], emax[0], &emin[1], &emax[1]);
    });
}

int qdx_ham_grid(const qdx_ham *h, qdx_grid *out) {
    return guarded([&] {
        if (!h || !out) fail(QDX_ERR_PARAM, "NULL argument");
        *out = qdx_grid{};
        out->ndims = h->grid.nd;
        for (int d = 0; d < h->grid.nd; ++d) {
            out->N[d] = h->grid.N[d];
            out->xmin[d] = h->grid.xmin[d];
            out->xmax[d] = h->grid.xmax[d];
        }
    });
}

static void ham_apply_impl(qdx_ham *h, qdx_wf *dst, qdx_wf *src, long step) {
    if (!h || !h->ctx) fail(QDX_ERR_PARAM, "Hamiltonian has no device context");
    check_wf(dst);
    check_wf(src);
    if (!(src->grid == h->grid) || src->S != h->S || !(dst->grid == h->grid) || dst->S != h->S)
        fail(QDX_ERR_INCOMPAT, "wave function does not match the Hamiltonian");
    if (dst == src) fail(QDX_ERR_PARAM, "in-place apply is not supported");
    h->finalize();
    HamDev<double> hd = h->dev.h64;
    h->set_fields(hd, step);
    LinePlan plans[kMaxDims];
    for (int d = 0; d < h->grid.nd; ++d) plans[d] = h->ctx->plan(h->grid.N[d], false);
    if (h->grid.nd > 1 && !h->X64.p) h->X64.alloc((size_t)h->S * h->grid.size() * 2 * sizeof(double));
    Launch<double>::ham_apply(h->ctx->stream, h->grid.dev(), hd, plans, src->ptr(), dst->ptr(), h->X64.as<double>());
}

int qdx_ham_apply(qdx_ham *h, qdx_wf *dst, qdx_wf *src, int64_t step) {
    return guarded([&] {
        ham_apply_impl(h, dst, src, step);
        HIPCHK(hipStreamSynchronize(h->ctx->stream));
    });
}

int qdx_ham_expec(qdx_ham *h, qdx_wf *psi, int64_t step, double *out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "out is NULL");
        check_wf(psi);
        qdx_wf tmp;
        tmp.ctx = psi->ctx;
        tmp.grid = psi->grid;
        tmp.S = psi->S;
        tmp.d.alloc(psi->d.bytes);
        ham_apply_impl(h, &tmp, psi, step);
        std::vector<double> per;
        overlap_states(psi, &tmp, per);
        double re = 0;
        for (int s = 0; s < psi->S; ++s) re += per[2 * s];
        *out = re * psi->grid.weight();   // Operator.h:226-252
    });
}

int qdx_ham_expec_noalloc(qdx_ham *h, qdx_wf *psi, int64_t step, double *out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "out is NULL");
        if (!h || !h->ctx) fail(QDX_ERR_PARAM, "Hamiltonian has no device context");
        check_wf(psi);
        if (!(psi->grid == h->grid) || psi->S != h->S) fail(QDX_ERR_INCOMPAT, "wave function does not match the Hamiltonian");
        if (psi->ctx != h->ctx) fail(QDX_ERR_INCOMPAT, "wave function and Hamiltonian on different contexts");
        qdx_ctx *c = h->ctx;
        const long n = psi->n();
        const long pts = (long)psi->S * n;
        if (pts > kNoAllocPoints)
            fail(QDX_ERR_MEMORY, "qdx_ham_expec_noalloc: " + std::to_string(pts) + " grid points times states exceed its "
                                 "scratch of " + std::to_string(kNoAllocPoints));
        // set-up happens elsewhere: the operator arrays are uploaded and the
        // transform tables built by the first qdx_ham_apply / qdx_ham_expec
        if (!h->final_) fail(QDX_ERR_STATE, "qdx_ham_expec_noalloc: apply the Hamiltonian once first (set-up)");
        LinePlan plans[kMaxDims];
        for (int d = 0; d < h->grid.nd; ++d) {
            if (!c->tw64.count(h->grid.N[d]))
                fail(QDX_ERR_STATE, "qdx_ham_expec_noalloc: apply the Hamiltonian once first (set-up)");
            plans[d] = c->plan(h->grid.N[d], false);   // cached table: no allocation
        }
        unsigned char *base = static_cast<unsigned char *>(qdx::noalloc_scratch(c->device));
        if (!base) fail(QDX_ERR_STATE, "device scratch was not set up for this device");
        // layout: H psi | X | partials | per-state sums
        double *hpsi = reinterpret_cast<double *>(base);
        double *X = reinterpret_cast<double *>(base + (size_t)kNoAllocPoints * 16);
        double *partials = reinterpret_cast<double *>(base + (size_t)2 * kNoAllocPoints * 16);
        double *scal = partials + kMaxPartials;
        HamDev<double> hd = h->dev.h64;
        h->set_fields(hd, step);
        Launch<double>::ham_apply(c->stream, h->grid.dev(), hd, plans, psi->ptr(), hpsi, X);
        launch_overlap(c->stream, psi->ptr(), hpsi, n, psi->S, partials, scal);
        HIPCHK(hipMemcpyAsync(c->pinned, scal, 2 * psi->S * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double re = 0;
        for (int s = 0; s < psi->S; ++s) re += c->pinned[2 * s];
        *out = re * psi->grid.weight();
    });
}

int qdx_ham_destroy(qdx_ham *h) {
    return guarded([&] {
        if (!h) return;
        qdx_ctx *c = h->ctx;
        if (c) (void)hipStreamSynchronize(c->stream);
        delete h;
        ctx_release(c);
    });
}

int qdx_bessel_j0(int32_t n, double x, double *out, int32_t *nz) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "out is NULL");
        std::vector<double> b;
        int z = bessel_j0(n, x, b);
        std::copy(b.begin(), b.end(), out);
        if (nz) *nz = z;
    });
}

int qdx_cheby_params(qdx_ham *h, double dt, int32_t order, double prec, double scaling, qdx_cheby_info *info,
                     double *coeff, int32_t coeff_len) {
    return guarded([&] {
        if (!h || !info) fail(QDX_ERR_PARAM, "NULL argument");
        ChebyParams c = cheby_params(h, dt, order, prec, scaling);
        info->order = c.order;
        info->nz = c.nz;
        info->offset = c.offset;
        info->scaling = c.scaling;
        info->prec = c.prec;
        info->emin = c.emin;
        info->emax = c.emax;
        info->offset_i = c.offset_i;
        if (coeff)
            for (int i = 0; i < std::min<int>(coeff_len, c.order); ++i) coeff[i] = c.coeff[i];
    });
}

static void cheby_create_impl(qdx_ham *h, double dt, int32_t order, double prec, double scaling, int32_t precision,
                              bool imag, qdx_prop **out) {
        if (!h || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (!h->ctx) fail(QDX_ERR_PARAM, "Hamiltonian has no device context");
        if (precision != QDX_FP64 && precision != QDX_FP32_BULK && precision != QDX_COMPACT)
            fail(QDX_ERR_PARAM, "unknown precision");
        // the imaginary-time series cancels down to exp(-R dt) of its terms:
        // FP32 terms would leave nothing of the result
        if (imag && precision != QDX_FP64) fail(QDX_ERR_PARAM, "imaginary time needs FP64");
        auto p = std::make_unique<qdx_prop>();
        p->kind = qdx_prop::CHEBY;
        p->ctx = h->ctx;
        p->precision = precision;
        p->dt = dt;
        p->grid = h->grid;
        p->S = h->S;
        p->ham = h;
        p->imag = imag;
        p->cp = cheby_params(h, dt, order, prec, scaling, imag);
        // compact terms on grids with nd >= 2 and one or two states; other
        // grids keep their terms in registers or take the FP64 kernels
        p->cmp = precision == QDX_COMPACT && h->grid.nd >= 2 && h->S <= 2;
        if (precision == QDX_FP32_BULK || p->cmp) fp32_series(p->cp, dt);
        // imaginary time: every order up to the series length (its break
        // test compares consecutive orders, not the coefficients)
        p->k_launch = imag ? p->cp.order : launch_window(p->cp, dt);
        if (precision == QDX_FP32_BULK) h->ensure32();
        // element bytes of the term buffers (compact: 8 hi + 2 lo)
        const size_t es = p->cmp ? 10 : precision == QDX_FP32_BULK ? 2 * sizeof(float) : 2 * sizeof(double);
        const size_t nbytes = (size_t)p->S * p->grid.size() * es;
        p->phiA.alloc(nbytes);
        p->phiB.alloc(nbytes);
        if (p->grid.nd > 1) p->X.alloc(nbytes);
        p->ctl.alloc(sizeof(ChebyCtl));
        HIPCHK(hipMemset(p->ctl.p, 0, sizeof(ChebyCtl)));
        p->partials.alloc(4 * kMaxPartials * sizeof(double));
        // (FP64 lines above 2048 points need more registers than a wave of a
        // 512-thread block holds; they keep the per-order kernels, and so
        // does imaginary time)
        // the whole-step kernel keeps its line buffers, the twiddle table
        // and the 2 * k_launch coefficients in LDS (160 KiB per CU)
        const size_t es_line = precision == QDX_FP32_BULK ? 8 : 16;
        const size_t L0 = (size_t)p->grid.N[0];
        const size_t line_lds = (2 * (L0 + L0 / 8) + (L0 + L0 / 8)) * es_line + 16 * (size_t)p->k_launch;
        p->line_step = !imag && p->grid.nd == 1 && p->S == 1 &&
                       p->grid.N[0] <= (precision == QDX_FP32_BULK ? kMaxLine : 2048) &&
                       line_lds <= 160 * 1024;
        if (p->line_step) {
            // the same products as the per-order scalars of cheby_step
            const std::complex<double> ph =
                std::exp(std::complex<double>(0, -1) * dt * std::complex<double>(p->cp.offset, p->cp.offset_i));
            std::vector<double> c(2 * (size_t)std::max(p->k_launch, 1));
            for (int k = 0; k < p->k_launch; ++k) {
                const std::complex<double> v = p->cp.coeff[k] * ph;
                c[2 * k] = v.real();
                c[2 * k + 1] = v.imag();
            }
            p->coefd.alloc(c.size() * sizeof(double));
            HIPCHK(hipMemcpy(p->coefd.p, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        HIPCHK(hipDeviceSynchronize());
        ctx_acquire(p->ctx);
        *out = p.release();
}

int qdx_cheby_create(qdx_ham *h, double dt, int32_t order, double prec, double scaling, int32_t precision,
                     qdx_prop **out) {
    return guarded([&] { cheby_create_impl(h, dt, order, prec, scaling, precision, false, out); });
}

int qdx_cheby_create_imag(qdx_ham *h, double dt, int32_t order, double prec, double scaling, qdx_prop **out) {
    return guarded([&] { cheby_create_impl(h, dt, order, prec, scaling, QDX_FP64, true, out); });
}

int qdx_bessel_i0(int32_t n, double x, double *out, int32_t *nz) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "out is NULL");
        std::vector<double> b;
        int z = bessel_i0_scaled(n, x, b);
        std::copy(b.begin(), b.end(), out);
        if (nz) *nz = z;
    });
}

int qdx_cheby_info_get(qdx_prop *p, qdx_cheby_info *info) {
    return guarded([&] {
        if (!p || !info || p->kind != qdx_prop::CHEBY) fail(QDX_ERR_PARAM, "not a Chebyshev propagator");
        info->order = p->cp.order;
        info->nz = p->cp.nz;
        info->offset = p->cp.offset;
        info->scaling = p->cp.scaling;
        info->prec = p->cp.prec;
        info->emin = p->cp.emin;
        info->emax = p->cp.emax;
        info->offset_i = p->cp.offset_i;
    });
}

int qdx_cheby_coeffs(qdx_prop *p, double *coeff) {
    return guarded([&] {
        if (!p || !coeff || p->kind != qdx_prop::CHEBY) fail(QDX_ERR_PARAM, "not a Chebyshev propagator");
        std::copy(p->cp.coeff.begin(), p->cp.coeff.end(), coeff);
    });
}

static void gspo_create_impl(qdx_ham *const *ops, int32_t nops, double dt, int32_t precision, bool imag,
                             qdx_prop **out) {
    {
        if (!ops || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (nops < 1) fail(QDX_ERR_PARAM, "GSPO is empty!");
        if (dt == 0) fail(QDX_ERR_PARAM, "No time step defined");
        if (precision != QDX_FP64 && precision != QDX_FP32_BULK && precision != QDX_COMPACT)
            fail(QDX_ERR_PARAM, "unknown precision");
        // QDX_COMPACT has no GSPO terms to store: psi stays complex128 (FP64)
        if (precision == QDX_COMPACT) precision = QDX_FP64;
        auto p = std::make_unique<qdx_prop>();
        p->kind = qdx_prop::GSPO;
        p->precision = precision;
        p->dt = dt;
        p->imag = imag;
        for (int i = 0; i < nops; ++i) {
            qdx_ham *o = ops[i];
            if (!o || !o->ctx) fail(QDX_ERR_PARAM, "GSPO misses an operator. Sequence must be A, B, C, ...");
            o->finalize();
            if (o->has_nip()) fail(QDX_ERR_INCOMPAT, "GSPO: a negative imaginary potential has no split operator here");
            if (i == 0) {
                p->ctx = o->ctx;
                p->grid = o->grid;
                p->S = o->S;
            } else if (!(o->grid == p->grid) || o->S != p->S || o->ctx != p->ctx) {
                fail(QDX_ERR_INCOMPAT, "GSPO operators differ in grid or states");
            }
            // ODSpace requirement: one operator kind per split operator
            const TermKind k0 = o->terms[0].kind;
            for (auto &t : o->terms) {
                if ((t.kind == T_KIN) != (k0 == T_KIN))
                    fail(QDX_ERR_INCOMPAT, "GSPO: Operator has no diagonal representation");
            }
            bool coupled = false, has_dip = false;
            for (auto &t : o->terms) {
                if (t.row != t.col) coupled = true;
                if (t.kind == T_DIP) has_dip = true;
            }
            if (o->multistate && has_dip)
                fail(QDX_ERR_INCOMPAT, "GSPO: multistate dipole split operators are not supported");
            if (coupled && o->S > 2) fail(QDX_ERR_INCOMPAT, "GSPO: coupled split operators need S <= 2");
            if (precision == QDX_FP32_BULK) o->ensure32();
            p->ops.push_back(o);
        }
        const size_t es = precision == QDX_FP32_BULK ? 2 * sizeof(float) : 2 * sizeof(double);
        const size_t nbytes = (size_t)p->S * p->grid.size() * es;
        if (p->grid.nd > 1) p->X.alloc(nbytes);
        if (precision == QDX_FP32_BULK) p->psi32.alloc(nbytes);
        p->partials.alloc(kMaxPartials * sizeof(double));
        ctx_acquire(p->ctx);
        *out = p.release();
    }
}

int qdx_gspo_create(qdx_ham *const *ops, int32_t nops, double dt, int32_t precision, qdx_prop **out) {
    return guarded([&] { gspo_create_impl(ops, nops, dt, precision, false, out); });
}

int qdx_gspo_create_imag(qdx_ham *const *ops, int32_t nops, double dt, qdx_prop **out) {
    return guarded([&] { gspo_create_impl(ops, nops, dt, QDX_FP64, true, out); });
}

int qdx_prop_apply(qdx_prop *p, qdx_wf *psi, int64_t step, qdx_step_stats *stats) {
    return guarded([&] {
        if (!p) fail(QDX_ERR_PARAM, "propagator is NULL");
        check_prop_wf(p, psi);
        hipStream_t s = p->ctx->stream;
        if (p->kind == qdx_prop::GSPO && p->precision == QDX_FP32_BULK)
            launch_convert_c128_to_c64(s, psi->ptr(), p->psi32.as<float>(), (long)p->S * psi->n());
        prop_step(p, psi, step);
        if (p->kind == qdx_prop::GSPO && p->precision == QDX_FP32_BULK)
            launch_convert_c64_to_c128(s, p->psi32.as<float>(), psi->ptr(), (long)p->S * psi->n());
        if (stats) {
            if (p->kind == qdx_prop::CHEBY) {
                ChebyCtl c;
                HIPCHK(hipMemcpyAsync(&c, p->ctl.p, sizeof(c), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                stats->n_hpsi = c.last_order;
                stats->nconv = cheby_last_nconv(p);
            } else if (p->kind == qdx_prop::SIL || p->kind == qdx_prop::ARNOLDI) {
                stats->n_hpsi = p->sil_napply;
                stats->nconv = 0;
            } else {
                stats->n_hpsi = 1;
                stats->nconv = 0;
            }
            std::vector<double> per;
            overlap_states(psi, psi, per);
            double tot = 0;
            for (int k = 0; k < psi->S; ++k) tot += per[2 * k];
            stats->norm2 = tot * psi->grid.weight();
        } else {
            HIPCHK(hipStreamSynchronize(s));
        }
    });
}

int qdx_prop_run(qdx_prop *p, qdx_wf *psi, int64_t first_step, int64_t nsteps, qdx_wf *psi0, double *norm2,
                 double *autocorr, int32_t *n_hpsi_per_step, qdx_step_stats *stats) {
    return guarded([&] {
        if (!p) fail(QDX_ERR_PARAM, "propagator is NULL");
        check_prop_wf(p, psi);
        if (nsteps < 0) fail(QDX_ERR_PARAM, "negative step count");
        if (autocorr && !psi0) fail(QDX_ERR_PARAM, "autocorrelation needs psi0");
        if (psi0) {
            check_wf(psi0);
            if (!(psi0->grid == psi->grid) || psi0->S != psi->S) fail(QDX_ERR_INCOMPAT, "psi0 does not match psi");
        }
        hipStream_t s = p->ctx->stream;
        const int S = p->S;
        const long n = psi->n();
        const bool g32 = p->kind == qdx_prop::GSPO && p->precision == QDX_FP32_BULK;
        // device observables: per step, per state (re, im) of <psi|psi> and <psi|psi0>
        const size_t per_step = 4 * S;
        p->obs.alloc(std::max<size_t>(1, (size_t)nsteps * per_step + nsteps) * sizeof(double));
        double *obs = p->obs.as<double>();
        int *orders = reinterpret_cast<int *>(obs + nsteps * per_step);
        if (g32) {
            launch_convert_c128_to_c64(s, psi->ptr(), p->psi32.as<float>(), (long)S * n);
            if (autocorr) {
                p->psi0_32.alloc((size_t)S * n * 2 * sizeof(float));
                launch_convert_c128_to_c64(s, psi0->ptr(), p->psi0_32.as<float>(), (long)S * n);
            }
        }
        double *part = p->partials.as<double>();
        if (!part || p->partials.bytes < kMaxPartials * sizeof(double))
            fail(QDX_ERR_STATE, "propagator has no reduction buffer for the observables");
        std::vector<int> sil_counts;
        const bool fused = nsteps > 0 && !p->backward && gspo_fusable(p).ok;
        if (fused) {
            // observables before step 0 here; those before the later steps
            // from the row passes between the steps (psi0: autocorrelation)
            double *o = obs;
            const bool want = norm2 || autocorr;
            if (g32) {
                float *ps = p->psi32.as<float>();
                const float *p0 = autocorr ? p->psi0_32.as<float>() : ps;
                if (norm2) launch_overlap_c64(s, ps, ps, n, S, part, o);
                if (autocorr) launch_overlap_c64(s, ps, p0, n, S, part, o + 2 * S);
                gspo_run_fused<float>(p, ps, p0, first_step, nsteps, want ? obs : nullptr, per_step);
            } else {
                double *ps = psi->ptr();
                const double *p0 = autocorr ? psi0->ptr() : ps;
                if (norm2) launch_overlap(s, ps, ps, n, S, part, o);
                if (autocorr) launch_overlap(s, ps, p0, n, S, part, o + 2 * S);
                gspo_run_fused<double>(p, ps, p0, first_step, nsteps, want ? obs : nullptr, per_step);
            }
        }
        for (int64_t i = 0; i < (fused ? 0 : nsteps); ++i) {
            double *o = obs + i * per_step;
            if (norm2) {
                if (g32) launch_overlap_c64(s, p->psi32.as<float>(), p->psi32.as<float>(), n, S, part, o);
                else launch_overlap(s, psi->ptr(), psi->ptr(), n, S, part, o);
            }
            if (autocorr) {
                if (g32) launch_overlap_c64(s, p->psi32.as<float>(), p->psi0_32.as<float>(), n, S, part, o + 2 * S);
                else launch_overlap(s, psi->ptr(), psi0->ptr(), n, S, part, o + 2 * S);
            }
            prop_step(p, psi, first_step + i);
            if (p->kind == qdx_prop::CHEBY)
                HIPCHK(hipMemcpyAsync(orders + i, &p->ctl.as<ChebyCtl>()->last_order, sizeof(int),
                                      hipMemcpyDeviceToDevice, s));
            if (p->kind == qdx_prop::SIL || p->kind == qdx_prop::ARNOLDI) sil_counts.push_back(p->sil_napply);
        }
        if (g32) launch_convert_c64_to_c128(s, p->psi32.as<float>(), psi->ptr(), (long)S * n);
        std::vector<double> host((size_t)nsteps * per_step + nsteps);
        if (nsteps > 0)
            HIPCHK(hipMemcpyAsync(host.data(), obs, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const double w = psi->grid.weight();
        int64_t total = 0;
        const int *hord = reinterpret_cast<const int *>(host.data() + nsteps * per_step);
        for (int64_t i = 0; i < nsteps; ++i) {
            const double *o = host.data() + i * per_step;
            if (norm2) {
                double t = 0;
                for (int k = 0; k < S; ++k) t += o[2 * k];
                norm2[i] = t * w;
                if (!std::isfinite(norm2[i])) fail(QDX_ERR_OVERFLOW, "Norm is not a number");
            }
            if (autocorr) {
                double re = 0, im = 0;
                for (int k = 0; k < S; ++k) {
                    re += o[2 * S + 2 * k];
                    im += o[2 * S + 2 * k + 1];
                }
                autocorr[2 * i] = re * w;
                autocorr[2 * i + 1] = im * w;
            }
            const int cnt = p->kind == qdx_prop::CHEBY                                   ? hord[i]
                            : (p->kind == qdx_prop::SIL || p->kind == qdx_prop::ARNOLDI) ? sil_counts[i]
                                                                                         : 1;
            if (n_hpsi_per_step) n_hpsi_per_step[i] = cnt;
            total += cnt;
        }
        if (stats) {
            stats->n_hpsi = total;
            stats->nconv = 0;
            if (p->kind == qdx_prop::CHEBY) stats->nconv = cheby_last_nconv(p);
            std::vector<double> per;
            overlap_states(psi, psi, per);
            double t = 0;
            for (int k = 0; k < S; ++k) t += per[2 * k];
            stats->norm2 = t * w;
        }
    });
}

// ---------------------------------------------------------------------------
// eigenstates by imaginary-time relaxation (ProgEigen::ImagTimeEF + DiagBasis)
// and the short iterative Lanczos propagator (OSIL)

extern "C++" {   // helpers with C++ types inside the extern "C" block
namespace {

// cyclic Jacobi rotations with accumulated eigenvectors; returns eigenvalues
// ascending and V (column i = eigenvector i, row-major n x n)
void sym_eigen_vectors(int n, std::vector<double> a, std::vector<double> &ev, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    double scale = 0;
    for (double x : a) scale += x * x;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
        if (off <= 1e-32 * scale) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[(size_t)p * n + q];
                if (apq == 0) continue;
                const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - sn * akq;
                    a[(size_t)k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - sn * aqk;
                    a[(size_t)q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - sn * vkq;
                    V[(size_t)k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
    ev.resize(n);
    std::vector<double> W((size_t)n * n);
    for (int i = 0; i < n; ++i) {
        ev[i] = a[(size_t)idx[i] * n + idx[i]];
        for (int k = 0; k < n; ++k) W[(size_t)k * n + i] = V[(size_t)k * n + idx[i]];
    }
    V.swap(W);
}

void chk(int rc) {
    if (rc != QDX_OK) fail(rc, g_err);
}

// <a|b> (qdx_wf_overlap); an out-parameter, since this file's C-linkage block
// holds it
void ovl(qdx_wf *a, qdx_wf *b, std::complex<double> &z) {
    double o[2];
    chk(qdx_wf_overlap(a, b, o));
    z = {o[0], o[1]};
}

// exp(A) e_0 for a small complex matrix A (n x n, row-major): scaling and
// squaring with the [13/13] Pade approximant (N. J. Higham, "The scaling and
// squaring method for the matrix exponential revisited", SIAM J. Matrix
// Anal. Appl. 26 (2005) 1179-1193), the linear system by LU with partial
// pivoting
void expm_col0(int n, const std::vector<std::complex<double>> &Ain, std::vector<std::complex<double>> &col0) {
    using cd = std::complex<double>;
    static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                                 1187353796428800.0,  129060195264000.0,   10559470521600.0,
                                 670442572800.0,      33522128640.0,       1323241920.0,
                                 40840800.0,          960960.0,            16380.0,
                                 182.0,               1.0};
    const double theta13 = 5.371920351148152;
    auto idx = [n](int i, int j) { return (size_t)i * n + j; };
    auto mul = [&](const std::vector<cd> &X, const std::vector<cd> &Y) {
        std::vector<cd> Z((size_t)n * n, 0.0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                const cd x = X[idx(i, k)];
                if (x == cd(0)) continue;
                for (int j = 0; j < n; ++j) Z[idx(i, j)] += x * Y[idx(k, j)];
            }
        return Z;
    };
    double nrm = 0;   // 1-norm
    for (int j = 0; j < n; ++j) {
        double c = 0;
        for (int i = 0; i < n; ++i) c += std::abs(Ain[idx(i, j)]);
        nrm = std::max(nrm, c);
    }
    int sq = 0;
    if (nrm > theta13) sq = (int)std::ceil(std::log2(nrm / theta13));
    std::vector<cd> A = Ain;
    const double f = std::ldexp(1.0, -sq);
    for (auto &a : A) a *= f;
    const std::vector<cd> A2 = mul(A, A), A4 = mul(A2, A2), A6 = mul(A4, A2);
    std::vector<cd> U1((size_t)n * n), V((size_t)n * n), W1((size_t)n * n), W2((size_t)n * n);
    for (size_t k = 0; k < A.size(); ++k) {
        W1[k] = b[13] * A6[k] + b[11] * A4[k] + b[9] * A2[k];
        W2[k] = b[12] * A6[k] + b[10] * A4[k] + b[8] * A2[k];
    }
    std::vector<cd> T1 = mul(A6, W1), T2 = mul(A6, W2);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const size_t k = idx(i, j);
            const double I = i == j ? 1.0 : 0.0;
            U1[k] = T1[k] + b[7] * A6[k] + b[5] * A4[k] + b[3] * A2[k] + b[1] * I;
            V[k] = T2[k] + b[6] * A6[k] + b[4] * A4[k] + b[2] * A2[k] + b[0] * I;
        }
    const std::vector<cd> U = mul(A, U1);
    // (V - U) X = (V + U)
    std::vector<cd> P((size_t)n * n), Q((size_t)n * n);
    for (size_t k = 0; k < P.size(); ++k) {
        P[k] = V[k] - U[k];
        Q[k] = V[k] + U[k];
    }
    std::vector<int> piv(n);
    for (int c = 0; c < n; ++c) {
        int pr = c;
        for (int r = c + 1; r < n; ++r)
            if (std::abs(P[idx(r, c)]) > std::abs(P[idx(pr, c)])) pr = r;
        if (P[idx(pr, c)] == cd(0)) fail(QDX_ERR_OVERFLOW, "matrix exponential: singular Pade denominator");
        if (pr != c)
            for (int j = 0; j < n; ++j) {
                std::swap(P[idx(c, j)], P[idx(pr, j)]);
                std::swap(Q[idx(c, j)], Q[idx(pr, j)]);
            }
        for (int r = c + 1; r < n; ++r) {
            const cd m = P[idx(r, c)] / P[idx(c, c)];
            if (m == cd(0)) continue;
            for (int j = c; j < n; ++j) P[idx(r, j)] -= m * P[idx(c, j)];
            for (int j = 0; j < n; ++j) Q[idx(r, j)] -= m * Q[idx(c, j)];
        }
    }
    std::vector<cd> X((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int r = n - 1; r >= 0; --r) {
            cd t = Q[idx(r, j)];
            for (int k = r + 1; k < n; ++k) t -= P[idx(r, k)] * X[idx(k, j)];
            X[idx(r, j)] = t / P[idx(r, r)];
        }
    for (int i = 0; i < sq; ++i) X = mul(X, X);
    col0.resize(n);
    for (int i = 0; i < n; ++i) col0[i] = X[idx(i, 0)];
}

struct WfHold {   // owns wave functions until released to the caller
    std::vector<qdx_wf *> w;
    ~WfHold() {
        for (qdx_wf *x : w)
            if (x) qdx_wf_destroy(x);
    }
    qdx_wf *make(const qdx_wf *like) {
        qdx_grid g{};
        g.ndims = like->grid.nd;
        for (int d = 0; d < like->grid.nd; ++d) {
            g.N[d] = like->grid.N[d];
            g.xmin[d] = like->grid.xmin[d];
            g.xmax[d] = like->grid.xmax[d];
        }
        qdx_wf *x = nullptr;
        chk(qdx_wf_create(like->ctx, &g, like->S, &x));
        w.push_back(x);
        return x;
    }
    void drop(qdx_wf *x) {   // destroy one held function now
        for (auto &y : w)
            if (y == x) {
                qdx_wf_destroy(y);
                y = nullptr;
            }
    }
    void release(const std::vector<qdx_wf *> &out) {   // ownership passes to the caller
        for (auto &y : w)
            for (qdx_wf *x : out)
                if (y == x) y = nullptr;
    }
};

// ProgEigen::DiagBasis (ProgEigen.cpp:93-138): diagonalise H in the basis ef
// (raw energies on the diagonal) and return the mixed functions
// psi_i = sum_j V(j, i) ef_j with the eigenvalues in `energies`; the raw set
// is destroyed.
std::vector<qdx_wf *> diag_basis(qdx_ham *H, const qdx_wf *like, const std::vector<qdx_wf *> &ef,
                                 const std::vector<double> &e_raw, WfHold &keep, WfHold &tmp, double *energies) {
    const int nef = (int)ef.size();
    qdx_wf *hbuf = tmp.make(like);
    std::vector<double> S((size_t)nef * nef, 0.0);
    for (int i = 0; i < nef; ++i) {
        S[(size_t)i * nef + i] = e_raw[i];
        for (int j = 0; j < i; ++j) {   // MatrixElement(bra = ef_i, ket = ef_j)
            chk(qdx_ham_apply(H, hbuf, ef[j], 0));
            std::complex<double> z;
            ovl(ef[i], hbuf, z);
            const double m = z.real();
            S[(size_t)j * nef + i] = S[(size_t)i * nef + j] = m;
        }
    }
    std::vector<double> ev, V;
    sym_eigen_vectors(nef, S, ev, V);
    std::vector<qdx_wf *> mixed;
    for (int i = 0; i < nef; ++i) {
        qdx_wf *m = keep.make(like);
        chk(qdx_wf_copy(m, ef[0]));
        launch_scale_axpy(m->ctx->stream, m->ptr(), nullptr, (long)m->S * m->n(), V[(size_t)0 * nef + i], 0, 0);
        for (int j = 1; j < nef; ++j) chk(qdx_wf_axpy(m, V[(size_t)j * nef + i], 0.0, ef[j]));
        mixed.push_back(m);
        energies[i] = ev[i];
    }
    for (qdx_wf *x : ef) keep.drop(x);
    return mixed;
}

// PeakFinder::Find(data) (tools/PeakFinder.cpp:83-126): the maximum of every
// region above mean + tol * sigma; a region still open at the end counts
// when the last value is not its maximum
std::vector<int> peaks_regions(const std::vector<double> &d, double tol) {
    const size_t n = d.size();
    double mean = 0, sig = 0;
    for (double v : d) mean += v;
    mean /= double(n);
    for (double v : d) sig += (v - mean) * (v - mean);
    sig = std::sqrt(sig / double(n));
    const double lim = mean + tol * sig;
    std::vector<int> pk;
    bool open = false;
    double dmax = 0;
    int imax = -1;
    for (size_t i = 0; i < n; ++i) {
        if (!open && d[i] > lim) {
            open = true;
            dmax = 0;
            imax = -1;
        }
        if (open && d[i] < lim) {
            open = false;
            pk.push_back(imax);
        }
        if (open && d[i] > dmax) {
            dmax = d[i];
            imax = (int)i;
        }
    }
    if (open && imax != -1 && d[n - 1] < dmax) pk.push_back(imax);
    return pk;
}

// PeakFinder::Find(data, winsize) (tools/PeakFinder.cpp:131-170, Maximas
// 177-190): every strict local maximum (d falls after rising) above
// mean + tol * sigma of the window [i - w, i + w) clipped to the data
std::vector<int> peaks_window(const std::vector<double> &d, double tol, int w) {
    const int n = (int)d.size();
    std::vector<int> maxi;
    double delta = 0;
    for (int i = 1; i < n; ++i) {
        if (d[i] < d[i - 1] && delta > 0) maxi.push_back(i - 1);
        delta = d[i] - d[i - 1];
    }
    std::vector<int> pk;
    for (int m : maxi) {
        const int lo = m < w ? 0 : m - w;
        const int hi = m + w > n - 1 ? n - 1 : m + w;
        double mean = 0, sig = 0;
        for (int i = lo; i < hi; ++i) mean += d[i];
        mean /= double(hi - lo);
        for (int i = lo; i < hi; ++i) sig += (d[i] - mean) * (d[i] - mean);
        sig = std::sqrt(sig / double(hi - lo));
        if (d[m] > mean + tol * sig) pk.push_back(m);
    }
    return pk;
}

// forward DFT, unnormalised, e^{-2 pi i k t / n} (FFTW_FORWARD): radix 2 when
// n is a power of two, else the direct sum
void dft_forward(std::vector<std::complex<double>> &a) {
    const size_t n = a.size();
    if (n && (n & (n - 1)) == 0) {
        for (size_t i = 1, j = 0; i < n; ++i) {   // bit reversal
            size_t bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) std::swap(a[i], a[j]);
        }
        for (size_t len = 2; len <= n; len <<= 1) {
            for (size_t i = 0; i < n; i += len)
                for (size_t k = 0; k < len / 2; ++k) {
                    const double ang = -2.0 * M_PI * double(k) / double(len);
                    const std::complex<double> w(std::cos(ang), std::sin(ang));
                    const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                    a[i + k] = u + v;
                    a[i + k + len / 2] = u - v;
                }
        }
        return;
    }
    std::vector<std::complex<double>> out(n);
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> acc = 0;
        for (size_t t = 0; t < n; ++t) {
            const double ang = -2.0 * M_PI * double((k * t) % n) / double(n);
            acc += a[t] * std::complex<double>(std::cos(ang), std::sin(ang));
        }
        out[k] = acc;
    }
    a.swap(out);
}

}  // namespace
}  // extern "C++"

int qdx_eigen_imag(qdx_prop *U, qdx_ham *Hin, const qdx_wf *psi0, int32_t nef, int32_t max_steps, double conv,
                   int32_t diag, qdx_wf **efs, double *energies, double *energies_raw, int32_t *steps) {
    return guarded([&] {
        if (!U || !psi0 || !efs || !energies) fail(QDX_ERR_PARAM, "NULL argument");
        if (!U->imag) fail(QDX_ERR_PARAM, "an imaginary-time propagator is required");
        if (nef < 1) fail(QDX_ERR_PARAM, "Nef must be >= 1");
        if (max_steps < 1) fail(QDX_ERR_PARAM, "steps must be >= 1");
        qdx_wf *p0 = const_cast<qdx_wf *>(psi0);
        check_prop_wf(U, p0);
        // the energies' Hamiltonian: given, else the propagator's own
        // (OPropagator::Hamiltonian, ProgEigen.cpp:697); GSPO has none
        qdx_ham *H = Hin ? Hin : U->ham;
        if (!H) fail(QDX_ERR_PARAM, "a Hamiltonian is required for the energies");
        if (!(H->grid == p0->grid) || H->S != p0->S || H->ctx != p0->ctx)
            fail(QDX_ERR_INCOMPAT, "Hamiltonian and wave function differ");
        WfHold keep, tmp;
        std::vector<qdx_wf *> ef;
        std::vector<double> e_raw(nef);
        qdx_wf *psi_old = tmp.make(p0);
        for (int i = 0; i < nef; ++i) {   // ProgEigen.cpp:365-404
            qdx_wf *psi = keep.make(p0);
            chk(qdx_wf_copy(psi, p0));
            int s = 0;
            double dff = 1;
            while (s < max_steps && dff > conv) {
                chk(qdx_wf_copy(psi_old, psi));
                prop_step(U, psi, s);
                if (i > 0) {   // OProjection::Apply (OProjection.cpp:133-149), psi -= P psi
                    std::vector<std::complex<double>> c(i);
                    for (int j = 0; j < i; ++j) ovl(ef[j], psi, c[j]);
                    for (int j = 0; j < i; ++j) chk(qdx_wf_axpy(psi, -c[j].real(), -c[j].imag(), ef[j]));
                }
                chk(qdx_wf_normalize(psi));
                std::complex<double> z;
                ovl(psi_old, psi, z);
                dff = std::abs(1.0 - z);
                ++s;
            }
            ef.push_back(psi);
            chk(qdx_ham_expec(H, psi, s, &e_raw[i]));
            if (steps) steps[i] = s;
        }
        if (energies_raw) std::copy(e_raw.begin(), e_raw.end(), energies_raw);
        if (diag) ef = diag_basis(H, p0, ef, e_raw, keep, tmp, energies);   // DiagBasis
        else std::copy(e_raw.begin(), e_raw.end(), energies);
        chk(qdx_ctx_synchronize(U->ctx));
        for (int i = 0; i < nef; ++i) efs[i] = ef[i];
        keep.release(ef);
    });
}

int qdx_eigen_ac(qdx_prop *U, qdx_ham *Hin, const qdx_wf *psi0, int32_t nsteps, int32_t win, double tol,
                 int32_t width, int32_t diag, int32_t max_ef, qdx_wf **efs, double *energies, double *energies_raw,
                 int32_t *nef, double *autocorr) {
    return guarded([&] {
        if (!U || !psi0 || !efs || !energies || !nef) fail(QDX_ERR_PARAM, "NULL argument");
        if (U->imag) fail(QDX_ERR_PARAM, "the autocorrelation method needs a real-time propagator");
        if (nsteps < 2) fail(QDX_ERR_PARAM, "Not enough steps for autocorrelation");
        if (max_ef < 1) fail(QDX_ERR_PARAM, "max_ef must be >= 1");
        qdx_wf *p0 = const_cast<qdx_wf *>(psi0);
        check_prop_wf(U, p0);
        qdx_ham *H = Hin ? Hin : U->ham;
        if (!H) fail(QDX_ERR_PARAM, "a Hamiltonian is required for the energies");
        if (!(H->grid == p0->grid) || H->S != p0->S || H->ctx != p0->ctx)
            fail(QDX_ERR_INCOMPAT, "Hamiltonian and wave function differ");
        WfHold keep, tmp;
        // ProgEigen.cpp:436-470: propagate and record <psi(t)|psi0>.  QDng
        // keeps every step's wave function (WFBuffer, swapped to disk beyond
        // its memory cap); here the propagation runs twice instead: this pass
        // records the autocorrelation, the second one (below, after the peak
        // search) adds each step into the peak functions.  The propagation is
        // deterministic, so both passes see the same psi(t) and the device
        // holds n_ef + 2 wave functions instead of nsteps.
        const int N = nsteps;
        std::vector<std::complex<double>> ac(N);
        ac[0] = 1.0;
        const int arn_order0 = U->arn_order;   // Arnoldi carries its order between steps
        qdx_wf *psi = tmp.make(p0);
        chk(qdx_wf_copy(psi, p0));
        for (int i = 1; i < N; ++i) {
            prop_step(U, psi, i - 1);
            ovl(psi, p0, ac[i]);
            if (std::isnan(ac[i].real())) fail(QDX_ERR_OVERFLOW, "Correlation is not a number");
        }
        if (autocorr)
            for (int i = 0; i < N; ++i) {
                autocorr[2 * i] = ac[i].real();
                autocorr[2 * i + 1] = ac[i].imag();
            }
        // Hann window (math_functions.h:108-113), forward FFT, |.|, peaks
        if (win)
            for (int n = 0; n < N; ++n) ac[n] *= 0.5 * (1. - std::cos(2 * M_PI * double(n) / double(N - 1)));
        dft_forward(ac);
        std::vector<double> power(N);
        for (int i = 0; i < N; ++i) power[i] = std::abs(ac[i]);
        const std::vector<int> pk = width > 0 ? peaks_window(power, tol, width) : peaks_regions(power, tol);
        const int n_ef = (int)pk.size();
        if (n_ef > max_ef) fail(QDX_ERR_PARAM, "more peaks than max_ef");
        if (n_ef == 0) fail(QDX_ERR_PARAM, "no peaks found in the spectrum");
        // ProgEigen.cpp:562-596: integrate every peak's function out
        const double dt = U->dt;
        const double dw = 2 * M_PI / (dt * double(N));
        std::vector<qdx_wf *> ef;
        std::vector<double> e_raw(n_ef), ek(n_ef);
        for (int k = 0; k < n_ef; ++k) {
            const int idx = pk[k];
            ek[k] = idx < N / 2 ? double(idx) * dw : double(idx - N) * dw;
            qdx_wf *f = keep.make(p0);
            chk(qdx_wf_copy(f, p0));
            ef.push_back(f);
        }
        // second pass: f_k = psi0 + sum_s exp(i E_k s dt) psi(s dt), each
        // f_k summed in step order
        U->arn_order = arn_order0;
        chk(qdx_wf_copy(psi, p0));
        for (int s = 1; s < N; ++s) {
            prop_step(U, psi, s - 1);
            for (int k = 0; k < n_ef; ++k) {
                const double ph = ek[k] * double(s) * dt;
                chk(qdx_wf_axpy(ef[k], std::cos(ph), std::sin(ph), psi));
            }
        }
        for (int k = 0; k < n_ef; ++k) {
            chk(qdx_wf_normalize(ef[k]));
            chk(qdx_ham_expec(H, ef[k], 0, &e_raw[k]));
        }
        if (energies_raw) std::copy(e_raw.begin(), e_raw.end(), energies_raw);
        if (diag) ef = diag_basis(H, p0, ef, e_raw, keep, tmp, energies);
        else std::copy(e_raw.begin(), e_raw.end(), energies);
        chk(qdx_ctx_synchronize(U->ctx));
        for (int i = 0; i < n_ef; ++i) efs[i] = ef[i];
        *nef = n_ef;
        keep.release(ef);
    });
}

// SIL and Arnoldi: `order` Krylov vectors and one buffer as device wave
// functions
static void krylov_create(qdx_ham *h, double dt, int32_t order, double err, int32_t imag, bool arnoldi,
                          qdx_prop **out) {
    {
        if (!h || !out) fail(QDX_ERR_PARAM, "NULL argument");
        if (!h->ctx) fail(QDX_ERR_PARAM, "Hamiltonian has no device context");
        if (dt == 0) fail(QDX_ERR_PARAM, "No time step defined");
        if (order <= 0) order = 20;                        // OSIL.h:11, OArnoldi.h:11
        if (order < 3) fail(QDX_ERR_PARAM, arnoldi ? "Arnoldi: order doesn't make sense" : "SIL: order doesn't make sense");
        h->finalize();
        if (h->has_nip() && (!arnoldi || imag))
            fail(QDX_ERR_INCOMPAT, "negative imaginary potential: SIL (Lanczos) needs a Hermitian Hamiltonian; "
                                   "Arnoldi runs it in real time only");
        auto p = std::make_unique<qdx_prop>();
        p->kind = arnoldi ? qdx_prop::ARNOLDI : qdx_prop::SIL;
        p->ctx = h->ctx;
        p->precision = QDX_FP64;
        p->dt = dt;
        p->grid = h->grid;
        p->S = h->S;
        p->ham = h;
        p->imag = imag != 0;
        p->sil_order = order;
        p->sil_err = err > 0 ? err : 1e-12;                // OSIL.h:12
        p->arn_conv = err > 0 ? err : 1e-12;               // OArnoldi.h:13
        p->arn_order = order;
        qdx_grid g{};
        g.ndims = h->grid.nd;
        for (int d = 0; d < h->grid.nd; ++d) {
            g.N[d] = h->grid.N[d];
            g.xmin[d] = h->grid.xmin[d];
            g.xmax[d] = h->grid.xmax[d];
        }
        auto cleanup = [&] {
            for (qdx_wf *w : p->lzb) qdx_wf_destroy(w);
            p->lzb.clear();
            if (p->sil_buf) qdx_wf_destroy(p->sil_buf);
            p->sil_buf = nullptr;
        };
        for (int i = 0; i <= order; ++i) {
            qdx_wf *w = nullptr;
            const int rc = qdx_wf_create(h->ctx, &g, h->S, &w);
            if (rc != QDX_OK) {
                const std::string e = g_err;
                cleanup();
                fail(rc, e);
            }
            if (i < order) p->lzb.push_back(w);
            else p->sil_buf = w;
        }
        // reduction partials of qdx_prop_run's observables
        p->partials.alloc(kMaxPartials * sizeof(double));
        ctx_acquire(p->ctx);
        *out = p.release();
    }
}

int qdx_sil_create(qdx_ham *h, double dt, int32_t order, double err, int32_t imag, qdx_prop **out) {
    return guarded([&] { krylov_create(h, dt, order, err, imag, false, out); });
}

int qdx_arnoldi_create(qdx_ham *h, double dt, int32_t order, double conv, int32_t imag, qdx_prop **out) {
    return guarded([&] { krylov_create(h, dt, order, conv, imag, true, out); });
}

int qdx_prop_set_direction(qdx_prop *p, int32_t backward) {
    return guarded([&] {
        if (!p) fail(QDX_ERR_PARAM, "propagator is NULL");
        if (p->imag) {
            if (!backward) fail(QDX_ERR_INCOMPAT, "imaginary-time propagators run backward (exponent -dt) only");
            return;
        }
        // backward = conj(forward(conj psi)) holds for self-conjugate H only
        if (backward && p->ham && p->ham->has_nip())
            fail(QDX_ERR_INCOMPAT, "backward propagation with a negative imaginary potential is not supported");
        p->backward = backward != 0;
    });
}

int qdx_prop_exponent(qdx_prop *p, double out[2]) {
    return guarded([&] {
        if (!p || !out) fail(QDX_ERR_PARAM, "NULL argument");
        // OPropagator::Exponent (OPropagator.h:44-155): -i dt forward, +i dt
        // backward; imaginary time backward -dt
        out[0] = p->imag ? -p->dt : 0.0;
        out[1] = p->imag ? 0.0 : (p->backward ? p->dt : -p->dt);
    });
}

int qdx_prop_destroy(qdx_prop *p) {
    return guarded([&] {
        if (!p) return;
        qdx_ctx *c = p->ctx;
        if (c) (void)hipStreamSynchronize(c->stream);
        for (qdx_wf *w : p->lzb) qdx_wf_destroy(w);
        if (p->sil_buf) qdx_wf_destroy(p->sil_buf);
        delete p;
        ctx_release(c);
    });
}

}  // extern "C"

namespace {

// OSIL::BuildLZB + DiagLZB (OSIL.cpp:79-190) over device operations: the
// Lanczos vectors live on the device, the tridiagonal matrix is diagonalised
// on the host.  QDng's own details are kept: q_0 is psi unnormalised, the
// error estimate dt^(it-2) b_0 prod_{i=1}^{it-3} b_i / i ends the iteration,
// the last diagonal element is recomputed as <q|H|q>, and the tridiagonal
// matrix keeps its full order (entries not reached stay 0).
void sil_step(qdx_prop *p, qdx_wf *psi, long step) {
    const int order = p->sil_order;
    std::vector<qdx_wf *> &q = p->lzb;
    qdx_wf *buf = p->sil_buf;
    qdx_ham *H = p->ham;
    hipStream_t s = p->ctx->stream;
    const long n = (long)p->S * psi->n();
    std::vector<double> alpha(order, 0.0), beta(order, 0.0);
    int napply = 0;
    auto norm_of = [&](qdx_wf *w) {
        double nn = 0;
        chk(qdx_wf_norm(w, &nn, nullptr));
        return nn;
    };
    auto scale = [&](qdx_wf *w, double f) { launch_scale_axpy(s, w->ptr(), nullptr, n, f, 0.0, 0.0); };
    std::complex<double> z;
    const double psinorm = norm_of(psi);
    chk(qdx_wf_copy(q[0], psi));
    chk(qdx_ham_apply(H, buf, q[0], step));
    ++napply;
    ovl(q[0], buf, z);
    alpha[0] = z.real();
    chk(qdx_wf_copy(q[1], buf));
    chk(qdx_wf_axpy(q[1], -alpha[0], 0.0, q[0]));
    beta[0] = norm_of(q[1]);
    scale(q[1], 1 / beta[0]);
    int convorder = 2, it;
    for (it = 2; it < order; it++) {
        double err = std::pow(p->dt, it - 2) * beta[0];
        for (int i = 1; i < it - 2; i++) err *= beta[i] / double(i);
        err = std::fabs(err);
        convorder++;
        chk(qdx_ham_apply(H, q[it], q[it - 1], step));
        ++napply;
        ovl(q[it - 1], q[it], z);
        alpha[it - 1] = z.real();
        chk(qdx_wf_axpy(q[it], -alpha[it - 1], 0.0, q[it - 1]));
        chk(qdx_wf_axpy(q[it], -beta[it - 2], 0.0, q[it - 2]));
        beta[it - 1] = norm_of(q[it]);
        scale(q[it], 1 / beta[it - 1]);
        if (err < p->sil_err * psinorm) break;
    }
    chk(qdx_ham_expec(H, q[it - 1], step, &alpha[it - 1]));
    ++napply;
    // Z e^{c Lambda} Z^T e_0, c = -i dt (real time) or -dt (imaginary)
    std::vector<double> T((size_t)order * order, 0.0), ev, Z;
    for (int i = 0; i < order; ++i) {
        T[(size_t)i * order + i] = alpha[i];
        if (i + 1 < order) T[(size_t)i * order + i + 1] = T[(size_t)(i + 1) * order + i] = beta[i];
    }
    sym_eigen_vectors(order, T, ev, Z);
    const std::complex<double> expo = p->imag ? std::complex<double>(-p->dt, 0) : std::complex<double>(0, -p->dt);
    std::vector<std::complex<double>> vect(order, 0.0);
    for (int k = 0; k < order; ++k) {
        const std::complex<double> w = std::exp(expo * ev[k]) * Z[(size_t)0 * order + k];
        for (int i = 0; i < order; ++i) vect[i] += Z[(size_t)i * order + k] * w;
    }
    // psi = vect_0 psi + sum_{i < convorder} vect_i q_i (q_0 holds psi)
    launch_scale_axpy(s, psi->ptr(), q[0]->ptr(), n, 0.0, vect[0].real(), vect[0].imag());
    for (int i = 1; i < convorder; ++i) chk(qdx_wf_axpy(psi, vect[i].real(), vect[i].imag(), q[i]));
    p->sil_napply = napply;
}

// OArnoldi::BuildLZB + DiagLZB + Propagate (OArnoldi.cpp:64-190) over device
// operations: modified Gram-Schmidt against all previous vectors, the
// Hessenberg matrix's last column from plain products with H q, and the
// combination exp(c H_A) e_0 (here by Pade instead of QDng's
// eigendecomposition ZR e^{c Lambda} ZR^-1: the same matrix function).
// The converged order carries over to the next step, as in OArnoldi.
void arnoldi_step(qdx_prop *p, qdx_wf *psi, long step) {
    const int M = p->sil_order;
    std::vector<qdx_wf *> &q = p->lzb;
    qdx_wf *buf = p->sil_buf;
    qdx_ham *H = p->ham;
    hipStream_t s = p->ctx->stream;
    const long n = (long)p->S * psi->n();
    std::vector<std::complex<double>> tHA((size_t)M * M, 0.0);
    int napply = 0;
    std::complex<double> z;
    double psinorm = 0;
    chk(qdx_wf_norm(psi, &psinorm, nullptr));
    chk(qdx_wf_copy(q[0], psi));
    for (int it = 1; it < M; it++) {
        chk(qdx_ham_apply(H, buf, q[it - 1], step));
        ++napply;
        for (int j = 0; j <= it - 1; j++) {
            ovl(q[j], buf, z);
            tHA[(size_t)j * M + it - 1] = z;
            chk(qdx_wf_axpy(buf, -z.real(), -z.imag(), q[j]));
        }
        double nb = 0;
        chk(qdx_wf_norm(buf, &nb, nullptr));
        if (nb < 1e-11) break;   // ARNOLDI_UNDERRUN (OArnoldi.h:12)
        tHA[(size_t)it * M + it - 1] = nb;
        chk(qdx_wf_copy(q[it], buf));
        launch_scale_axpy(s, q[it]->ptr(), nullptr, n, 1. / nb, 0.0, 0.0);
        p->arn_order = it + 1;
        std::complex<double> err = std::pow(p->dt, it - 2) * tHA[(size_t)1 * M + 0] * psinorm;
        for (int i = 1; i < it - 2; i++) err *= tHA[(size_t)(i + 1) * M + i] / double(i);
        if (std::abs(err) < p->arn_conv) break;
    }
    const int order = p->arn_order;
    chk(qdx_ham_apply(H, buf, q[order - 1], step));
    ++napply;
    for (int j = 0; j < order; j++) {
        ovl(q[j], buf, z);
        tHA[(size_t)j * M + order - 1] = z;
    }
    const std::complex<double> expo = p->imag ? std::complex<double>(-p->dt, 0) : std::complex<double>(0, -p->dt);
    std::vector<std::complex<double>> A((size_t)order * order), vect;
    for (int i = 0; i < order; ++i)
        for (int j = 0; j < order; ++j) A[(size_t)i * order + j] = expo * tHA[(size_t)i * M + j];
    expm_col0(order, A, vect);
    launch_scale_axpy(s, psi->ptr(), q[0]->ptr(), n, 0.0, vect[0].real(), vect[0].imag());
    for (int i = 1; i < order; ++i) chk(qdx_wf_axpy(psi, vect[i].real(), vect[i].imag(), q[i]));
    p->sil_napply = napply;
}

}  // namespace
