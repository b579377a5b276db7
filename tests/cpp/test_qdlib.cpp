// Tests of the QDng-style C++ classes (include/qdx_qdlib.hpp) on the GPU,
// written after QDng's CppUnit tests with grid operators in place of the
// level-system ones:
//   TestCheby (src/tests/TestCheby.cpp:60-125): parameters after Init and a
//     two-state system whose upper state turns by -i over dt = pi;
//   OGSPOTest (src/tests/OGSPOTest.cpp:96-121): a coupling of 1/2 moves the
//     whole population over dt = pi (half a Rabi cycle);
//   OGridNablaSqTest (src/tests/OGridNablaSqTest.cpp:76-137): the periodic
//     sine is an eigenfunction with eigenvalue 2 (pi / (L + dx))^2 / (2 m);
// plus a harmonic-oscillator propagation (Chebyshev against GSPO, norm) and
// QDng's exceptions for bad parameters.
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "qdx_qdlib.hpp"

using namespace qdxlib;

static int g_fail = 0, g_run = 0;
#define LOOSE_EPS 1e-10
#define ROUGH_EPS 1e-6

#define ASSERT_NEAR(exp, got, eps)                                                              \
    do {                                                                                        \
        const double e_ = (exp), g_ = (got);                                                    \
        if (!(std::fabs(e_ - g_) <= (eps))) {                                                   \
            std::printf("  FAIL %s:%d: expected %.17g got %.17g\n", __FILE__, __LINE__, e_, g_); \
            throw 1;                                                                            \
        }                                                                                       \
    } while (0)
#define ASSERT_TRUE(c)                                                          \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);          \
            throw 1;                                                            \
        }                                                                       \
    } while (0)
template <class E, class F>
static void assert_throws(F f, int line) {
    try {
        f();
    } catch (const E &) {
        return;
    }
    std::printf("  FAIL line %d: expected exception not thrown\n", line);
    throw 1;
}
#define ASSERT_THROW(E, stmt) assert_throws<E>([&] { stmt; }, __LINE__)

static void run(const char *name, const std::function<void()> &f) {
    ++g_run;
    std::printf("[cpp] %s\n", name);
    std::fflush(stdout);
    try {
        f();
    } catch (int) {
        ++g_fail;
    } catch (const std::exception &e) {
        std::printf("  FAIL: exception %s\n", e.what());
        ++g_fail;
    }
}

static qdx_grid line_grid(int N, double xmin, double xmax) {
    qdx_grid g{};
    g.ndims = 1;
    g.N[0] = N;
    g.xmin[0] = xmin;
    g.xmax[0] = xmax;
    return g;
}

// two states on 8 points, V00 = 0, V11 = 1/2 (TestCheby's matrix)
struct TwoLevel {
    OGridPotential v00, v11;
    OMultistate H;
    TwoLevel() {
        v00.SetData(std::vector<double>(8, 0.0));
        v11.SetData(std::vector<double>(8, 0.5));
        H.Add(&v00, 0, 0);
        H.Add(&v11, 1, 1);
    }
};

int main() {
    const qdx_grid g8 = line_grid(8, -1, 1);

    run("TestCheby::API_Test", [&] {
        TwoLevel sys;
        QDClock clock;
        clock.Dt(M_PI);
        clock.Steps(2);
        OCheby U;
        U.Clock(&clock);
        ASSERT_TRUE(U.Name() == "OCheby");
        U.Add("hamiltonian", &sys.H);
        WFMultistate psi0;
        psi0.Init(g8, 2);
        ParamContainer pm;
        U.Init(pm);
        U.Init(&psi0);
        double d;
        ASSERT_TRUE(U.Params().GetValue("prec", d));
        ASSERT_NEAR(1e-10, d, 1e-30);
        U.Params().GetValue("offset.imag", d);
        ASSERT_NEAR(0, d, 1e-30);
        U.Params().GetValue("offset.real", d);
        ASSERT_NEAR(0.25, d, 1e-30);
        U.Params().GetValue("scaling", d);
        ASSERT_NEAR(0.25, d, 1e-30);
    });

    run("TestCheby::NUMERIC_Test", [&] {
        TwoLevel sys;
        QDClock clock;
        clock.Dt(M_PI);
        OCheby U;
        U.Clock(&clock);
        U.Add("hamiltonian", &sys.H);
        WFMultistate psi0, psi1;
        psi0.Init(g8, 2);
        psi1.Init(g8, 2);
        for (int x = 0; x < 8; ++x) {
            psi0.at(0, x) = 1;
            psi0.at(1, x) = 0;
        }
        ParamContainer pm;
        U.Init(pm);
        U.Init(&psi0);
        // psi = (1, 0): nothing happens
        U.Apply(&psi1, &psi0);
        for (int x = 0; x < 8; ++x) {
            ASSERT_NEAR(1, psi1.at(0, x).real(), LOOSE_EPS);
            ASSERT_NEAR(0, psi1.at(0, x).imag(), LOOSE_EPS);
            ASSERT_NEAR(0, std::abs(psi1.at(1, x)), LOOSE_EPS);
        }
        // psi = (0, 1): exp(-i pi/2) = -i
        for (int x = 0; x < 8; ++x) {
            psi0.at(0, x) = 0;
            psi0.at(1, x) = 1;
        }
        U.Apply(&psi1, &psi0);
        for (int x = 0; x < 8; ++x) {
            ASSERT_NEAR(0, std::abs(psi1.at(0, x)), LOOSE_EPS);
            ASSERT_NEAR(0, psi1.at(1, x).real(), LOOSE_EPS);
            ASSERT_NEAR(-1, psi1.at(1, x).imag(), LOOSE_EPS);
        }
    });

    run("OGSPOTest::NUMERIC_Test_H0", [&] {
        OGridPotential c01;
        c01.SetData(std::vector<double>(8, 0.5));
        OMultistate W;
        W.Add(&c01, 0, 1);
        ParamContainer ms;
        ms.SetValue("unity", false);
        W.Init(ms);
        QDClock clock;
        clock.Dt(M_PI);
        OGSPO U;
        U.Clock(&clock);
        ASSERT_TRUE(U.Name() == "OGSPO");
        U.Add("A", &W);
        WFMultistate psi0, psi1;
        psi0.Init(g8, 2);
        psi1.Init(g8, 2);
        for (int x = 0; x < 8; ++x) psi0.at(0, x) = 1;
        ParamContainer pm;
        U.Init(pm);
        U.Init(&psi0);
        U.Apply(&psi1, &psi0);
        for (int x = 0; x < 8; ++x) {
            ASSERT_NEAR(0, std::norm(psi1.at(0, x)), LOOSE_EPS);
            ASSERT_NEAR(1, std::norm(psi1.at(1, x)), LOOSE_EPS);
        }
        psi1 = psi0;
        U.Apply(&psi1);
        for (int x = 0; x < 8; ++x) {
            ASSERT_NEAR(0, std::norm(psi1.at(0, x)), LOOSE_EPS);
            ASSERT_NEAR(1, std::norm(psi1.at(1, x)), LOOSE_EPS);
        }
    });

    for (int N : {64, 96}) {
        run(N == 64 ? "OGridNablaSqTest::NUMERIC_Test (N = 64)" : "OGridNablaSqTest::NUMERIC_Test (N = 96)", [&] {
            const double L = 10, dx = L / (N - 1);
            const double pre = 2 * (M_PI / (L + dx)) * (M_PI / (L + dx));
            ParamContainer gp;
            gp.SetValue("dims", 1);
            gp.SetValue("N0", N);
            gp.SetValue("xmin0", -5.0);
            gp.SetValue("xmax0", 5.0);
            WFGridCartesian wf;
            wf.Init(gp);
            std::unique_ptr<WFGridCartesian> wf2(wf.NewInstance());
            for (int i = 0; i < N; ++i) wf[i] = std::sin(2 * M_PI * i / N);   // period L + dx
            OGridNablaSq T;
            ASSERT_TRUE(T.Name() == "OGridNablaSq");
            ParamContainer p;
            p.SetValue("mass", 1.0);
            T.Init(p);
            OSum H;
            H.Add(&T);
            H.Apply(wf2.get(), &wf);
            for (int i = 0; i < N; ++i) {
                ASSERT_NEAR(pre * wf[i].real(), (*wf2)[i].real(), LOOSE_EPS);
                ASSERT_NEAR(pre * wf[i].imag(), (*wf2)[i].imag(), LOOSE_EPS);
            }
            // Emax of the kinetic operator: (pi^2 / 2) / (m dx^2)
            ASSERT_NEAR(M_PI * M_PI / 2 / (dx * dx), H.Emax().real(), 1e-9);
        });
    }

    run("harmonic oscillator: OCheby and OGSPO agree, norm and energy conserved", [&] {
        const int N = 256;
        ParamContainer gp;
        gp.SetValue("dims", 1);
        gp.SetValue("N0", N);
        gp.SetValue("xmin0", -8.0);
        gp.SetValue("xmax0", 8.0);
        WFGridCartesian a, b;
        a.Init(gp);
        b.Init(gp);
        const double dx = 16.0 / (N - 1);
        std::vector<double> V(N);
        for (int i = 0; i < N; ++i) {
            const double x = -8 + i * dx;
            V[i] = 0.05 * x * x;
            a[i] = b[i] = std::exp(-(x - 1) * (x - 1));
        }
        a.Normalize();
        b.Normalize();
        OGridNablaSq T;
        ParamContainer p;
        p.SetValue("mass", 1836.0);
        T.Init(p);
        OGridPotential pot;
        pot.SetData(V);
        OSum H;
        H.Add(&T);
        H.Add(&pot);
        OSum Tk;
        Tk.Add(&T);
        OSum Vp;
        Vp.Add(&pot);
        QDClock clock;
        clock.Dt(2.0);
        OCheby U;
        U.Clock(&clock);
        U.Add("hamiltonian", &H);
        OGSPO G;
        G.Clock(&clock);
        G.Add("A", &Tk);
        G.Add("B", &Vp);
        const double e0 = H.Expec(&a);
        for (int s = 0; s < 50; ++s) {
            clock.TimeStep(s);
            U.Apply(&a);
            G.Apply(&b);
        }
        // the series stops at a norm change below prec = 1e-10 per step
        ASSERT_NEAR(1.0, a.Norm(), 1e-8);
        ASSERT_NEAR(e0, H.Expec(&a), 1e-8);
        const double ov = std::abs(a * b);
        ASSERT_NEAR(1.0, ov, 1e-4);   // second-order splitting error of GSPO at dt = 2
    });

    // QDng code holds operators and propagators as Operator* / OPropagator*
    // (ProgPropa, Reporter, ProgEigen): the same calls through the virtual
    // interface (Operator.h:142-369, OPropagator.h:44-169)
    run("Operator interface through Operator*", [&] {
        const int N = 64;
        const qdx_grid g = line_grid(N, -6, 6);
        const double dx = 12.0 / (N - 1);
        WFMultistate psi;
        psi.Init(g, 1);
        std::vector<double> V(N);
        for (int i = 0; i < N; ++i) {
            const double x = -6 + i * dx;
            V[i] = 0.05 * x * x + 0.01;
            psi.at(0, i) = dcomplex(std::exp(-(x - 0.5) * (x - 0.5)), 0.3 * std::exp(-x * x));
        }
        OGridNablaSq T;
        ParamContainer pm;
        pm.SetValue("mass", 900.0);
        T.Init(pm);
        OGridPotential Vp;
        Vp.SetData(V);
        OSum H;
        H.Add(&T);
        H.Add(&Vp);
        std::vector<Operator *> ops = {&T, &Vp, &H};
        std::vector<std::unique_ptr<WaveFunction>> out;
        for (Operator *o : ops) {
            o->Init(&psi);
            std::unique_ptr<Operator> fresh(o->NewInstance());
            ASSERT_TRUE(fresh->Name() == o->Name());
            out.emplace_back(psi.NewInstance());
            o->Apply(out.back().get(), &psi);
            // in place equals out of place
            std::unique_ptr<WaveFunction> w(psi.NewInstance());
            *w = psi;
            o->Apply(w.get());
            for (int i = 0; i < N; ++i) ASSERT_NEAR(0, std::abs((*w)[i] - (*out.back())[i]), 1e-14);
            // Expec is Re <psi|O|psi> = Re MatrixElement(psi, psi)
            ASSERT_NEAR(o->MatrixElement(&psi, &psi).real(), o->Expec(&psi), 1e-12);
            ASSERT_TRUE(o->Hermitian());
        }
        // H psi = T psi + V psi through ApplyAdd (Operator.h:307)
        std::unique_ptr<WaveFunction> tv(psi.NewInstance());
        T.Apply(tv.get(), &psi);
        Vp.ApplyAdd(tv.get(), &psi);
        for (int i = 0; i < N; ++i) ASSERT_NEAR(0, std::abs((*tv)[i] - (*out[2])[i]), 1e-13);
        // bounds: kinetic (pi^2/2)/(m dx^2) and 0, potential max/min, sum of both
        ASSERT_NEAR(M_PI * M_PI / 2 / (900.0 * dx * dx), T.Emax().real(), 1e-12);
        ASSERT_NEAR(0.0, T.Emin().real(), 1e-15);
        ASSERT_NEAR(*std::max_element(V.begin(), V.end()), Vp.Emax().real(), 1e-15);
        ASSERT_NEAR(*std::min_element(V.begin(), V.end()), Vp.Emin().real(), 1e-15);
        ASSERT_NEAR(T.Emax().real() + Vp.Emax().real(), H.Emax().real(), 1e-12);
    });

    run("OPropagator interface: exponent, backward, imaginary time", [&] {
        const int N = 128;
        const qdx_grid g = line_grid(N, -8, 8);
        const double dx = 16.0 / (N - 1);
        std::vector<double> V(N);
        WFMultistate psi;
        psi.Init(g, 1);
        for (int i = 0; i < N; ++i) {
            const double x = -8 + i * dx;
            V[i] = 0.05 * x * x;
            psi.at(0, i) = std::exp(-(x - 1) * (x - 1));
        }
        psi.Normalize();
        OGridNablaSq T;
        ParamContainer pm;
        pm.SetValue("mass", 1836.0);
        T.Init(pm);
        OGridPotential Vp;
        Vp.SetData(V);
        OSum H, Tk, Vk;
        H.Add(&T);
        H.Add(&Vp);
        Tk.Add(&T);
        Vk.Add(&Vp);
        QDClock clock;
        clock.Dt(4.0);
        OCheby cheby;
        cheby.Add("hamiltonian", &H);
        OSIL sil;
        sil.Add("hamiltonian", &H);
        OArnoldi arn;
        arn.Add("hamiltonian", &H);
        OGSPO gspo;
        gspo.Add("A", &Tk);
        gspo.Add("B", &Vk);
        std::vector<OPropagator *> props = {&cheby, &sil, &arn, &gspo};
        for (OPropagator *U : props) {
            ASSERT_THROW(EIncompatible, U->Backward());     // no clock yet
            U->Clock(&clock);
            ASSERT_NEAR(-4.0, U->Exponent().imag(), 0);    // forward: -i dt
            ASSERT_NEAR(0.0, U->Exponent().real(), 0);
            std::unique_ptr<WaveFunction> w(psi.NewInstance());
            *w = psi;
            U->Apply(w.get());                              // Init on first use
            ASSERT_TRUE(U->Hamiltonian() != nullptr);
            const double e0 = U->Hamiltonian()->Expec(&psi), e1 = U->Hamiltonian()->Expec(w.get());
            ASSERT_NEAR(e0, e1, U == &gspo ? 1e-5 : 1e-9);  // GSPO: splitting error
            U->Backward();                                   // exponent +i dt
            ASSERT_NEAR(4.0, U->Exponent().imag(), 0);
            ASSERT_TRUE(U->Conj());
            U->Apply(w.get());                               // back to psi
            for (int i = 0; i < N; ++i) ASSERT_NEAR(0, std::abs((*w)[i] - psi.at(0, i)), 1e-9);
            U->Conj(false);                                  // Conj(false) is Forward()
            ASSERT_NEAR(-4.0, U->Exponent().imag(), 0);
            ASSERT_NEAR(0.0, U->Emax().real(), 0);           // propagators have no bounds
        }
        // imaginary time runs backward: exponent -dt, the energy decreases
        OCheby im;
        im.Add("hamiltonian", &H);
        im.Clock(&clock);
        im.ImaginaryTime();
        ASSERT_THROW(EIncompatible, im.Init(&psi));        // forward imaginary time: exp(+H dt)
        im.Backward();
        ASSERT_NEAR(-4.0, im.Exponent().real(), 0);
        std::unique_ptr<WaveFunction> w(psi.NewInstance());
        *w = psi;
        OPropagator *Ui = &im;
        Ui->Apply(w.get());
        w->Normalize();
        ASSERT_TRUE(H.Expec(w.get()) < H.Expec(&psi));
    });

    run("OGobbler: absorbing boundary (negative imaginary potential)", [&] {
        const int N = 128;
        const qdx_grid g = line_grid(N, -8, 8);
        const double dx = 16.0 / (N - 1);
        WFMultistate psi;
        psi.Init(g, 1);
        std::vector<double> V(N);
        for (int i = 0; i < N; ++i) {
            const double x = -8 + i * dx;
            V[i] = 0.002 * x * x;
            psi.at(0, i) = std::exp(-(x - 2) * (x - 2)) * dcomplex(std::cos(8 * x), std::sin(8 * x));
        }
        psi.Normalize();
        OGridNablaSq T;
        ParamContainer pm;
        pm.SetValue("mass", 1836.0);
        T.Init(pm);
        OGridPotential Vp;
        Vp.SetData(V);
        OGobbler nip, wall;
        ParamContainer pn, pw;
        pn.SetValue("rp0", 5.0);
        pn.SetValue("order", 6);
        pn.SetValue("gain", 0.05);
        pn.SetValue("nip", "true");
        nip.Init(pn);
        pw.SetValue("lp0", -6.0);
        pw.SetValue("gain", 0.2);
        pw.SetValue("residue", "true");
        wall.Init(pw);
        ASSERT_TRUE(!nip.Hermitian() && wall.Hermitian());
        ASSERT_NEAR(-0.05, nip.Emin().imag(), 0);            // OGobbler.cpp:188-192
        ASSERT_NEAR(0.2, wall.Emax().real(), 0);
        OSum H;
        H.Add(&T);
        H.Add(&Vp);
        H.Add(&nip);
        H.Add(&wall);
        H.Init(&psi);
        ASSERT_NEAR(-0.05, H.Emin().imag(), 0);              // the sum's complex Emin
        ASSERT_NEAR(0.0, H.Emax().imag(), 0);
        // residue: Expec = <O psi|O psi> (OGobbler.cpp:170-176)
        wall.Init(&psi);
        std::unique_ptr<WaveFunction> ow(psi.NewInstance());
        wall.Apply(ow.get(), &psi);
        ASSERT_NEAR((*ow * *ow).real(), wall.Expec(&psi), 1e-15);
        QDClock clock;
        clock.Dt(5.0);
        OCheby cheby;
        cheby.Add("hamiltonian", &H);
        cheby.Clock(&clock);
        std::unique_ptr<WaveFunction> w(psi.NewInstance());
        *w = psi;
        double last = 1.0;
        for (int k = 0; k < 40; ++k) {   // the packet runs into the absorber
            cheby.Apply(w.get());
            const double n2 = (*w * *w).real();
            ASSERT_TRUE(std::sqrt(n2) <= last + 1e-12);
            last = std::sqrt(n2);
        }
        ASSERT_TRUE(last < 0.9);
        ASSERT_THROW(EIncompatible, cheby.Backward());      // H is not self-conjugate
        ASSERT_THROW(EIncompatible, nip.Conj(true));
        OSIL sil;                                            // Lanczos needs a Hermitian H
        sil.Add("hamiltonian", &H);
        sil.Clock(&clock);
        ASSERT_THROW(EIncompatible, sil.Apply(w.get()));
    });

    run("exceptions", [&] {
        ParamContainer p;
        p.SetValue("mass", 0.0);
        OGridNablaSq T;
        ASSERT_THROW(EParamProblem, T.Init(p));
        ParamContainer gp;
        gp.SetValue("dims", 1);
        gp.SetValue("N0", 100);   // not 8 * 2^a * 3^b
        gp.SetValue("xmin0", -1.0);
        gp.SetValue("xmax0", 1.0);
        WFGridCartesian wf;
        ASSERT_THROW(EParamProblem, wf.Init(gp));
        OCheby U;
        QDClock clock;
        clock.Dt(1.0);
        U.Clock(&clock);
        WFMultistate psi;
        psi.Init(line_grid(8, -1, 1), 2);
        ASSERT_THROW(EParamProblem, U.Init(&psi));          // no hamiltonian
        TwoLevel sys;
        OCheby U2;
        U2.Clock(&clock);
        U2.Add("hamiltonian", &sys.H);
        WFMultistate three;
        three.Init(line_grid(8, -1, 1), 3);
        ASSERT_THROW(EIncompatible, U2.Init(&three));       // 2x2 operator on 3 states
    });

    std::printf("[cpp] %d tests, %d failed\n", g_run, g_fail);
    return g_fail == 0 ? 0 : 1;
}
