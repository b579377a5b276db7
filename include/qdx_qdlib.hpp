// qdx_qdlib.hpp - QDng-style C++ classes over the qdx C ABI (header only).
//
// The classes carry QDng's names and calling conventions for the grid
// propagation path (src/qdlib of QDng): parameters through a ParamContainer,
// operators initialised with Init(ParamContainer&) and Init(WaveFunction*),
// Apply(dst, src) / Apply(psi), Expec, Emax / Emin, propagators bound to a
// QDClock, and QDng's exception types.  All numerics run in libqdx.so on the
// GPU; wave functions keep their data on the device and copy it to the host
// on element access.
//
//   QDng class (file)                            here
//   ParamContainer (tools/ParamContainer.h)      ParamContainer
//   QDClock (tools/QDClock.h)                    QDClock
//   WFGridCartesian (qdlib/WFGridCartesian.h)    WFGridCartesian
//   WFMultistate (qdlib/WFMultistate.h)          WFMultistate
//   OGridNablaSq (qdlib/OGridNablaSq.h)          OGridNablaSq     param "mass"
//   OGridPotential (qdlib/OGridPotential.h)      OGridPotential   data via SetData
//   OGridDipole (qdlib/OGridDipole.h)            OGridDipole      + laser samples
//   OSum (qdlib/OSum.h)                          OSum             Add(Operator*)
//   OMultistate (qdlib/OMultistate.h)            OMultistate      Add(Operator*, i, j), "unity"
//   Operator (qdlib/Operator.h)                  Operator         virtual interface
//   OPropagator (qdlib/OPropagator.h)            OPropagator      Exponent, Forward, Backward, ...
//   OCheby (qdlib/OCheby.h)                      OCheby           Add("hamiltonian", H)
//   OSIL, OArnoldi (qdlib/OSIL.h, OArnoldi.h)    OSIL, OArnoldi   Add("hamiltonian", H)
//   OGSPO (qdlib/OGSPO.h)                        OGSPO            Add("A"/"B"/..., op)
//   EParamProblem ... (tools/Exception.h)        same names
#ifndef QDX_QDLIB_HPP
#define QDX_QDLIB_HPP

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "qdx.h"

namespace qdxlib {

using dcomplex = std::complex<double>;
using std::string;

// ---------------------------------------------------------------------------
// exceptions (tools/Exception.h)
class Exception : public std::runtime_error {
  public:
    explicit Exception(const string &m) : std::runtime_error(m) {}
};
class EParamProblem : public Exception {
  public:
    explicit EParamProblem(const string &m) : Exception("Parameter problem: " + m) {}
};
class EIncompatible : public Exception {
  public:
    explicit EIncompatible(const string &m) : Exception("Incompatible: " + m) {}
};
class EOverflow : public Exception {
  public:
    explicit EOverflow(const string &m) : Exception("Overflow: " + m) {}
};
class EMemError : public Exception {
  public:
    explicit EMemError(const string &m) : Exception("Memory: " + m) {}
};

inline void check(int rc) {
    if (rc == QDX_OK) return;
    const string m = qdx_last_error();
    switch (rc) {
    case QDX_ERR_PARAM: throw EParamProblem(m);
    case QDX_ERR_INCOMPAT: throw EIncompatible(m);
    case QDX_ERR_OVERFLOW: throw EOverflow(m);
    case QDX_ERR_MEMORY: throw EMemError(m);
    default: throw Exception(m);
    }
}

// ---------------------------------------------------------------------------
// ParamContainer: string-valued parameters with typed access
class ParamContainer {
    std::map<string, string> p_;

  public:
    template <class T>
    void SetValue(const string &name, const T &v) {
        std::ostringstream o;
        o.precision(17);
        o << std::boolalpha << v;
        p_[name] = o.str();
    }
    void SetValue(const string &name, const char *v) { p_[name] = v; }
    void SetValue(const string &name, const string &v) { p_[name] = v; }
    bool isPresent(const string &name) const { return p_.count(name) != 0; }
    bool GetValue(const string &name, string &v) const {
        auto it = p_.find(name);
        if (it == p_.end()) return false;
        v = it->second;
        return true;
    }
    bool GetValue(const string &name, double &v) const {
        string s;
        if (!GetValue(name, s)) return false;
        v = std::stod(s);
        return true;
    }
    bool GetValue(const string &name, int &v) const {
        string s;
        if (!GetValue(name, s)) return false;
        v = std::stoi(s);
        return true;
    }
    bool GetValue(const string &name, bool &v, bool def = false) const {
        string s;
        if (!GetValue(name, s)) {
            v = def;
            return false;
        }
        v = (s == "true" || s == "1" || s == "yes");
        return true;
    }
    // comma-separated list, as ParamContainer::GetArray (ParamContainer.cpp)
    bool GetArray(const string &name, std::vector<double> &a) const {
        string s;
        if (!GetValue(name, s)) return false;
        a.clear();
        std::istringstream in(s);
        string tok;
        while (std::getline(in, tok, ',')) a.push_back(std::stod(tok));
        return true;
    }
};

// ---------------------------------------------------------------------------
// QDClock: time step dt and step counter
class QDClock {
    double dt_ = 0;
    int steps_ = 0, step_ = 0;

  public:
    void Dt(double dt) { dt_ = dt; }
    double Dt() const { return dt_; }
    void Steps(int n) { steps_ = n; }
    int Steps() const { return steps_; }
    int TimeStep() const { return step_; }
    void TimeStep(int s) { step_ = s; }
    void Begin() { step_ = 0; }
    QDClock &operator++() {
        ++step_;
        return *this;
    }
    double Time() const { return step_ * dt_; }
};

// One device context per process (device 0), shared by all objects.
inline qdx_ctx *context() {
    static struct Holder {
        qdx_ctx *c = nullptr;
        Holder() { check(qdx_ctx_create(0, &c)); }
        ~Holder() {
            if (c) qdx_ctx_destroy(c);
        }
    } h;
    return h.c;
}

// ---------------------------------------------------------------------------
// grid geometry (qdlib/GridSystem.h): dx = (xmax - xmin)/(N - 1)
class GridSystem {
  protected:
    qdx_grid g_{};

  public:
    int Dim() const { return g_.ndims; }
    int DimSize(int d) const { return g_.N[d]; }
    double Xmin(int d) const { return g_.xmin[d]; }
    double Xmax(int d) const { return g_.xmax[d]; }
    double Dx(int d) const { return (g_.xmax[d] - g_.xmin[d]) / (g_.N[d] - 1); }
    long Size() const {
        long n = 1;
        for (int d = 0; d < g_.ndims; ++d) n *= g_.N[d];
        return n;
    }
    const qdx_grid &grid() const { return g_; }
    // "dims", "N0".., "xmin0".., "xmax0".. (GridSystem.proto)
    void InitGrid(const ParamContainer &pm) {
        int nd = 0;
        if (!pm.GetValue("dims", nd) || nd < 1 || nd > 3) throw EParamProblem("No dimensions given");
        g_.ndims = nd;
        for (int d = 0; d < nd; ++d) {
            const string i = std::to_string(d);
            if (!pm.GetValue("N" + i, g_.N[d])) throw EParamProblem("Number of points missing");
            if (!pm.GetValue("xmin" + i, g_.xmin[d]) || !pm.GetValue("xmax" + i, g_.xmax[d]))
                throw EParamProblem("Grid boundaries missing");
        }
    }
};

// ---------------------------------------------------------------------------
// wave functions: device data, host copy on access
class WaveFunction {
  protected:
    qdx_wf *wf_ = nullptr;
    mutable std::vector<dcomplex> host_;
    mutable bool host_valid_ = false, host_dirty_ = false;
    int S_ = 1;
    long n_ = 0;
    qdx_grid g_{};

    void alloc(const qdx_grid &g, int S) {
        g_ = g;
        S_ = S;
        n_ = 1;
        for (int d = 0; d < g.ndims; ++d) n_ *= g.N[d];
        check(qdx_wf_create(context(), &g_, S_, &wf_));
        host_.assign((size_t)n_ * S_, dcomplex(0));
        host_valid_ = true;
        host_dirty_ = true;
    }

  public:
    WaveFunction() = default;
    WaveFunction(const WaveFunction &) = delete;
    virtual ~WaveFunction() {
        if (wf_) qdx_wf_destroy(wf_);
    }
    // a new vector of the same kind on the same grid, values zero
    // (WaveFunction::NewInstance, qdlib/WaveFunction.h)
    virtual WaveFunction *NewInstance() const = 0;
    long size() const { return n_ * S_; }
    int States() const { return S_; }
    // the device vector, current (host changes uploaded)
    qdx_wf *device() {
        if (host_dirty_) {
            for (int s = 0; s < S_; ++s)
                check(qdx_wf_upload(wf_, s, reinterpret_cast<const double *>(host_.data() + (size_t)s * n_)));
            host_dirty_ = false;
        }
        return wf_;
    }
    // call after a device operation changed the vector: the device copy is
    // the current one (pending host changes were uploaded by device())
    void DeviceChanged() {
        host_valid_ = false;
        host_dirty_ = false;
    }
    dcomplex &operator[](long i) {
        fetch();
        host_dirty_ = true;
        return host_[i];
    }
    dcomplex operator[](long i) const {
        const_cast<WaveFunction *>(this)->fetch();
        return host_[i];
    }
    void fetch() {
        if (host_valid_) return;
        for (int s = 0; s < S_; ++s)
            check(qdx_wf_download(wf_, s, reinterpret_cast<double *>(host_.data() + (size_t)s * n_)));
        host_valid_ = true;
    }
    // WaveFunction::Norm = sqrt(<psi|psi>) with the grid weights
    double Norm() {
        double n = 0;
        check(qdx_wf_norm(device(), &n, nullptr));
        return n;
    }
    void Normalize() {
        check(qdx_wf_normalize(device()));
        DeviceChanged();
    }
    // <this|b>
    dcomplex operator*(WaveFunction &b) {
        double o[2];
        check(qdx_wf_overlap(device(), b.device(), o));
        return dcomplex(o[0], o[1]);
    }
    WaveFunction &operator=(WaveFunction &o) {
        check(qdx_wf_copy(wf_, o.device()));
        DeviceChanged();
        return *this;
    }
    WaveFunction &operator=(dcomplex c) {
        fetch();
        for (auto &x : host_) x = c;
        host_dirty_ = true;
        return *this;
    }
    const qdx_grid &grid() const { return g_; }
};

class WFGridCartesian : public WaveFunction, public GridSystem {
  public:
    void Init(const ParamContainer &pm) {
        InitGrid(pm);
        alloc(GridSystem::g_, 1);
    }
    // a new vector on the same grid, values zero
    WFGridCartesian *NewInstance() const override {
        auto *w = new WFGridCartesian();
        w->GridSystem::g_ = GridSystem::g_;
        w->alloc(GridSystem::g_, 1);
        return w;
    }
};

// [S][N] on one grid (WFMultistate.cpp: states stored one after another)
class WFMultistate : public WaveFunction {
  public:
    void Init(const qdx_grid &g, int states) { alloc(g, states); }
    WFMultistate *NewInstance() const override {
        auto *w = new WFMultistate();
        w->alloc(g_, S_);
        return w;
    }
    // element x of state s
    dcomplex &at(int s, long x) { return (*this)[(long)s * n_ + x]; }
};

// ---------------------------------------------------------------------------
// operators: QDng's polymorphic interface (qdlib/Operator.h:60-374).  Code
// that holds an Operator* or an OPropagator* drives any of the classes below
// through these virtual methods.
class Operator {
  protected:
    ParamContainer params_;
    QDClock *clock_ = nullptr;
    bool conj_ = false;
    bool hermitian_ = true;
    long step() const { return clock_ ? clock_->TimeStep() : 0; }

  public:
    virtual ~Operator() = default;
    virtual Operator *NewInstance() const = 0;                      // Operator.h:142
    virtual const string &Name() const = 0;
    virtual void Init(ParamContainer &pm) { params_ = pm; }         // Operator.h:149
    virtual void Init(WaveFunction *psi) = 0;                        // Operator.h:161
    ParamContainer &Params() { return params_; }
    virtual void Clock(QDClock *c) { clock_ = c; }                   // Operator.h:190
    QDClock *Clock() const { return clock_; }
    virtual void UpdateTime() {}                                     // Operator.h:183
    bool Hermitian() const { return hermitian_; }                   // Operator.h:206
    bool Conj() const { return conj_; }
    virtual void Conj(bool c) { conj_ = c; }                         // Operator.h:219
    virtual dcomplex Emax() = 0;                                     // Operator.h:259
    virtual dcomplex Emin() = 0;                                     // Operator.h:266
    virtual void Apply(WaveFunction *dst, WaveFunction *src) = 0;    // Operator.h:276
    // in place (Operator.h:281): through a copy of psi
    virtual void Apply(WaveFunction *psi) {
        std::unique_ptr<WaveFunction> buf(psi->NewInstance());
        *buf = *psi;
        Apply(psi, buf.get());
    }
    // dst += O src (Operator.h:307)
    virtual void ApplyAdd(WaveFunction *dst, WaveFunction *src) {
        std::unique_ptr<WaveFunction> buf(src->NewInstance());
        Apply(buf.get(), src);
        check(qdx_wf_axpy(dst->device(), 1.0, 0.0, buf->device()));
        dst->DeviceChanged();
    }
    // <bra|O|ket> (Operator.h:226)
    virtual dcomplex MatrixElement(WaveFunction *bra, WaveFunction *ket) {
        std::unique_ptr<WaveFunction> buf(ket->NewInstance());
        Apply(buf.get(), ket);
        return *bra * *buf;
    }
    // Re <psi|O|psi> (Operator.h:249)
    virtual double Expec(WaveFunction *psi) { return MatrixElement(psi, psi).real(); }
    virtual bool Valid(WaveFunction *psi) { return psi != nullptr; }   // Operator.h:361
};

// operators held on the device as one qdx Hamiltonian, built on the first
// Init(WaveFunction*) from the terms the class contributes
class DeviceOperator : public Operator {
  protected:
    qdx_ham *ham_ = nullptr;
    qdx_grid g_{};
    virtual bool multistate() const { return false; }
    virtual bool unity() const { return true; }
    virtual int states() const { return 1; }
    void need() const {
        if (!ham_) throw EParamProblem(Name() + " used before Init(WaveFunction*)");
    }

  public:
    ~DeviceOperator() override {
        if (ham_) qdx_ham_destroy(ham_);
    }
    // the operator's terms at matrix element (row, col) of h (a single
    // operator at (0, 0); composites place their children)
    virtual void AddTerms(qdx_ham *h, int row, int col) const = 0;
    using Operator::Init;
    void Init(WaveFunction *psi) override {
        if (!psi) throw EParamProblem("Invalid WaveFunction");
        if (ham_) return;
        g_ = psi->grid();
        if (psi->States() != states()) throw EIncompatible("operator and wave function differ in the number of states");
        check(qdx_ham_create(context(), &g_, states(), multistate() ? 1 : 0, unity() ? 1 : 0, &ham_));
        AddTerms(ham_, 0, 0);
    }
    qdx_ham *ham() const { return ham_; }
    dcomplex Emax() override {
        need();
        double lo[2], hi[2];
        check(qdx_ham_bounds_c(ham_, lo, hi));
        return dcomplex(hi[0], hi[1]);
    }
    dcomplex Emin() override {
        need();
        double lo[2], hi[2];
        check(qdx_ham_bounds_c(ham_, lo, hi));
        return dcomplex(lo[0], lo[1]);
    }
    void Apply(WaveFunction *dst, WaveFunction *src) override {
        Init(src);
        check(qdx_ham_apply(ham_, dst->device(), src->device(), step()));
        dst->DeviceChanged();
    }
    using Operator::Apply;
    double Expec(WaveFunction *psi) override {
        Init(psi);
        double e = 0;
        check(qdx_ham_expec(ham_, psi->device(), step(), &e));
        return e;
    }
};

// grid operators (single-state leaves); in a composite they add their term at
// the element they are placed on
class OGridNablaSq : public DeviceOperator {
    std::vector<double> mass_;

  public:
    OGridNablaSq *NewInstance() const override { return new OGridNablaSq(); }
    const string &Name() const override {
        static const string n = "OGridNablaSq";
        return n;
    }
    // "mass": comma-separated masses per dimension (<= 0: no kinetic term
    // in that dimension), OGridNablaSq.cpp:28-58
    void Init(ParamContainer &pm) override {
        Operator::Init(pm);
        if (!pm.GetArray("mass", mass_) || mass_.empty()) throw EParamProblem("No mass defined");
        for (double m : mass_)
            if (m == 0) throw EParamProblem("Zero mass defined");
    }
    using DeviceOperator::Init;
    void AddTerms(qdx_ham *h, int row, int col) const override {
        if (row != col) throw EIncompatible("kinetic operator on an off-diagonal element");
        if (mass_.empty()) throw EParamProblem("No mass defined");
        std::vector<double> m(mass_);
        m.resize(4, -1.0);
        check(qdx_ham_add_kinetic(h, row, m.data()));
    }
};

class OGridPotential : public DeviceOperator {
  protected:
    std::vector<double> v_;

  public:
    OGridPotential *NewInstance() const override { return new OGridPotential(); }
    const string &Name() const override {
        static const string n = "OGridPotential";
        return n;
    }
    void SetData(const std::vector<double> &v) { v_ = v; }
    double &operator[](long i) { return v_[i]; }
    void AddTerms(qdx_ham *h, int row, int col) const override {
        if (v_.empty()) throw EParamProblem("GridPotential without data");
        check(qdx_ham_add_potential(h, row, col, v_.data()));
    }
};

// -mu(x) E(t) with E from the laser samples at the clock's step
class OGridDipole : public OGridPotential {
    std::vector<double> laser_;
    double scale_ = 1.0;

  public:
    OGridDipole *NewInstance() const override { return new OGridDipole(); }
    const string &Name() const override {
        static const string n = "OGridDipole";
        return n;
    }
    std::vector<double> &GetLaser() { return laser_; }
    void Scale(double s) { scale_ = s; }
    void AddTerms(qdx_ham *h, int row, int col) const override {
        if (v_.empty() || laser_.empty()) throw EParamProblem("GridDipole without dipole or laser");
        check(qdx_ham_add_dipole(h, row, col, v_.data(), laser_.data(), (int32_t)laser_.size(), scale_));
    }
};

// OGobbler (qdlib/OGobbler.cpp): Butterworth low / high passes with edges
// "lp<d>" / "rp<d>" (x units), "order" (20), "gain" (1); "nip" turns it into
// the negative imaginary potential -i v (non-Hermitian; single-state Sum,
// real-time Chebyshev and Arnoldi only); "residue": Expec = <O psi|O psi>
class OGobbler : public DeviceOperator {
    int order_ = 20;
    double gain_ = 1.0;
    bool nip_ = false, residue_ = false;
    int dims_ = 0;

  public:
    OGobbler *NewInstance() const override { return new OGobbler(); }
    const string &Name() const override {
        static const string n = "OGobbler";
        return n;
    }
    void Init(ParamContainer &pm) override {   // OGobbler.cpp:85-122
        Operator::Init(pm);
        if (pm.isPresent("order")) {
            pm.GetValue("order", order_);
            if (order_ < 1) throw EParamProblem("Filter order to small");
        }
        if (pm.isPresent("dims")) pm.GetValue("dims", dims_);
        pm.GetValue("nip", nip_, false);
        hermitian_ = !nip_;
        if (pm.isPresent("gain")) pm.GetValue("gain", gain_);
        pm.GetValue("residue", residue_, false);
    }
    using DeviceOperator::Init;
    void Conj(bool c) override {
        if (c && nip_) throw EIncompatible("backward propagation with a negative imaginary potential");
        Operator::Conj(c);
    }
    using Operator::Conj;
    dcomplex Emax() override { return dcomplex(gain_, 0); }                          // OGobbler.cpp:183-186
    dcomplex Emin() override { return nip_ ? dcomplex(0, -gain_) : dcomplex(0, 0); }  // OGobbler.cpp:188-192
    void AddTerms(qdx_ham *h, int row, int col) const override {
        qdx_grid g{};
        check(qdx_ham_grid(h, &g));
        if (dims_ != 0 && dims_ != g.ndims) throw EIncompatible("OGobbler: Number of dims doesn't match WF");
        double lp[4], rp[4];
        for (int d = 0; d < 4; ++d) {
            lp[d] = rp[d] = std::nan("");
            const string k = std::to_string(d);
            if (d < g.ndims && params_.isPresent("lp" + k)) params_.GetValue("lp" + k, lp[d]);
            if (d < g.ndims && params_.isPresent("rp" + k)) params_.GetValue("rp" + k, rp[d]);
        }
        check(qdx_ham_add_gobbler(h, row, col, lp, rp, order_, gain_, nip_ ? 1 : 0));
    }
    double Expec(WaveFunction *psi) override {   // OGobbler.cpp:166-181
        if (!residue_) return DeviceOperator::Expec(psi);
        std::unique_ptr<WaveFunction> ket(psi->NewInstance());
        Apply(ket.get(), psi);
        return (*ket * *ket).real();
    }
};

// OSum (qdlib/OSum.cpp): the sum of its operators on one state
class OSum : public DeviceOperator {
    std::vector<const DeviceOperator *> ops_;

  public:
    OSum *NewInstance() const override { return new OSum(); }
    const string &Name() const override {
        static const string n = "OSum";
        return n;
    }
    void Add(Operator *o) {
        auto *d = dynamic_cast<const DeviceOperator *>(o);
        if (!d) throw EIncompatible("OSum takes grid operators");
        ops_.push_back(d);
    }
    int Size() const { return (int)ops_.size(); }
    void AddTerms(qdx_ham *h, int row, int col) const override {
        if (ops_.empty()) throw EParamProblem("Sum Operator is empty");
        for (auto *o : ops_) o->AddTerms(h, row, col);
    }
};

// OMultistate (qdlib/OMultistate.cpp): operators at matrix elements (i, j);
// "unity" (default true) puts the identity on empty diagonal elements
class OMultistate : public DeviceOperator {
    struct E {
        const DeviceOperator *op;
        int i, j;
    };
    std::vector<E> elems_;
    bool multistate() const override { return true; }
    int states() const override {
        int n = 0;
        for (auto &e : elems_) n = std::max(n, std::max(e.i, e.j) + 1);
        return n;
    }
    bool unity() const override {
        bool u = true;
        params_.GetValue("unity", u, true);
        return u;
    }

  public:
    OMultistate *NewInstance() const override { return new OMultistate(); }
    const string &Name() const override {
        static const string n = "OMultistate";
        return n;
    }
    void Add(Operator *o, int i, int j) {
        auto *d = dynamic_cast<const DeviceOperator *>(o);
        if (!d) throw EIncompatible("OMultistate takes grid operators");
        elems_.push_back({d, i, j});
    }
    void AddTerms(qdx_ham *h, int row, int col) const override {
        if (elems_.empty()) throw EParamProblem("Multistate operator is empty");
        if (row != 0 || col != 0) throw EIncompatible("a Multistate operator inside another one");
        for (auto &e : elems_) e.op->AddTerms(h, e.i, e.j);
    }
};

// ---------------------------------------------------------------------------
// propagators (qdlib/OPropagator.h:23-186): exponent c of exp(c H) per step,
// -i dt forward and +i dt backward in real time; imaginary time runs
// backward (exponent -dt, as ProgEigen sets it up)
class OPropagator : public Operator {
  protected:
    qdx_prop *prop_ = nullptr;
    bool forward_ = true, imaginary_ = false;
    Operator *H_ = nullptr;
    void need_clock() const {
        if (!clock_) throw EIncompatible("Propagator has no clock");
    }
    void sync_direction() {
        if (prop_ && !imaginary_) check(qdx_prop_set_direction(prop_, forward_ ? 0 : 1));
    }
    // QDng's init guards (OPropagator.h:171-184)
    void check_init(WaveFunction *psi) const {
        if (!clock_) throw EParamProblem("Progagator is missing a clock");
        if (clock_->Dt() == 0) throw EParamProblem("No time step defined");
        if (!psi) throw EParamProblem("Invalid WaveFunction");
        if (imaginary_ && forward_) throw EIncompatible("imaginary time runs backward (exponent -dt) only");
    }

  public:
    ~OPropagator() override {
        if (prop_) qdx_prop_destroy(prop_);
    }
    void Forward() {
        need_clock();
        forward_ = true;
        conj_ = false;
        sync_direction();
    }
    void Backward() {
        need_clock();
        forward_ = false;
        conj_ = true;
        sync_direction();
    }
    void ImaginaryTime() {
        if (prop_) throw EIncompatible("set imaginary time before Init(WaveFunction*)");
        imaginary_ = true;
    }
    void RealTime() {
        if (prop_ && imaginary_) throw EIncompatible("set real time before Init(WaveFunction*)");
        imaginary_ = false;
    }
    void Conj(bool c) override { c ? Backward() : Forward(); }
    using Operator::Conj;
    // OPropagator::Exponent (OPropagator.h:148-155)
    dcomplex Exponent() const {
        need_clock();
        const double dt = clock_->Dt();
        if (imaginary_) return forward_ ? dt : -dt;
        return forward_ ? dcomplex(0, -dt) : dcomplex(0, dt);
    }
    bool ImaginaryTimeSet() const { return imaginary_; }
    Operator *Hamiltonian() { return H_; }
    // propagators have no spectral bounds (OPropagator.h:157-159)
    dcomplex Emax() override { return 0; }
    dcomplex Emin() override { return 0; }
    // one time step at the clock's step
    void Apply(WaveFunction *psi) override {
        if (!prop_) Init(psi);
        qdx_step_stats st;
        check(qdx_prop_apply(prop_, psi->device(), step(), &st));
        psi->DeviceChanged();
    }
    void Apply(WaveFunction *dst, WaveFunction *src) override {
        *dst = *src;
        Apply(dst);
    }
    using Operator::Init;
    qdx_prop *prop() const { return prop_; }
};

// the Hamiltonian of a propagator given as one device operator
inline DeviceOperator *as_device(Operator *o, const char *what) {
    auto *d = dynamic_cast<DeviceOperator *>(o);
    if (!d) throw EIncompatible(string(what) + ": the operator must be a grid operator or a Sum/Multistate of them");
    return d;
}

class OCheby : public OPropagator {
    int precision_ = QDX_FP64;

  public:
    OCheby *NewInstance() const override { return new OCheby(); }
    const string &Name() const override {
        static const string n = "OCheby";
        return n;
    }
    // Add("hamiltonian", H)
    void Add(const string &key, Operator *H) {
        if (key != "hamiltonian") throw EParamProblem("OCheby takes a hamiltonian");
        H_ = H;
    }
    void Precision(int p) { precision_ = p; }
    using Operator::Init;
    // params "order", "prec", "scaling" as OCheby::Init(ParamContainer&)
    void Init(WaveFunction *psi) override {
        if (!H_) throw EParamProblem("Chebychev propagator has no hamiltonian");
        check_init(psi);
        DeviceOperator *H = as_device(H_, "OCheby");
        H->Clock(clock_);
        H->Init(psi);
        int order = 0;
        double prec = -1, scaling = 0;
        params_.GetValue("order", order);
        params_.GetValue("prec", prec);
        params_.GetValue("scaling", scaling);
        if (prop_) qdx_prop_destroy(prop_);
        prop_ = nullptr;
        if (imaginary_) check(qdx_cheby_create_imag(H->ham(), clock_->Dt(), order, prec, scaling, &prop_));
        else check(qdx_cheby_create(H->ham(), clock_->Dt(), order, prec, scaling, precision_, &prop_));
        sync_direction();
        qdx_cheby_info info;
        check(qdx_cheby_info_get(prop_, &info));
        // the values OCheby::Init publishes (OCheby.cpp:443-493)
        params_.SetValue("order", info.order);
        params_.SetValue("prec", info.prec);
        params_.SetValue("scaling", info.scaling);
        params_.SetValue("offset.real", info.offset);
        params_.SetValue("offset.imag", 0.0);
    }
};

// Lanczos and Arnoldi: params "order" and "err" / "conv" (OSIL.cpp:79-190,
// OArnoldi.cpp:64-190)
class OSIL : public OPropagator {
  public:
    OSIL *NewInstance() const override { return new OSIL(); }
    const string &Name() const override {
        static const string n = "OSIL";
        return n;
    }
    void Add(const string &key, Operator *H) {
        if (key != "hamiltonian") throw EParamProblem("OSIL takes a hamiltonian");
        H_ = H;
    }
    using Operator::Init;
    void Init(WaveFunction *psi) override {
        if (!H_) throw EParamProblem("SIL propagator has no hamiltonian");
        check_init(psi);
        DeviceOperator *H = as_device(H_, "OSIL");
        H->Clock(clock_);
        H->Init(psi);
        int order = 0;
        double err = 0;
        params_.GetValue("order", order);
        params_.GetValue("err", err);
        if (prop_) qdx_prop_destroy(prop_);
        prop_ = nullptr;
        check(qdx_sil_create(H->ham(), clock_->Dt(), order, err, imaginary_ ? 1 : 0, &prop_));
        sync_direction();
    }
};

class OArnoldi : public OPropagator {
  public:
    OArnoldi *NewInstance() const override { return new OArnoldi(); }
    const string &Name() const override {
        static const string n = "OArnoldi";
        return n;
    }
    void Add(const string &key, Operator *H) {
        if (key != "hamiltonian") throw EParamProblem("OArnoldi takes a hamiltonian");
        H_ = H;
    }
    using Operator::Init;
    void Init(WaveFunction *psi) override {
        if (!H_) throw EParamProblem("Arnoldi propagator has no hamiltonian");
        check_init(psi);
        DeviceOperator *H = as_device(H_, "OArnoldi");
        H->Clock(clock_);
        H->Init(psi);
        int order = 0;
        double conv = 0;
        params_.GetValue("order", order);
        params_.GetValue("conv", conv);
        if (prop_) qdx_prop_destroy(prop_);
        prop_ = nullptr;
        check(qdx_arnoldi_create(H->ham(), clock_->Dt(), order, conv, imaginary_ ? 1 : 0, &prop_));
        sync_direction();
    }
};

class OGSPO : public OPropagator {
    std::map<string, Operator *> ops_;
    int precision_ = QDX_FP64;
    // H = A + B + ... of the split operators (OGSPO.cpp:61-105)
    class SplitSum : public DeviceOperator {
        std::vector<const DeviceOperator *> ops_;
        int S_ = 1;
        bool ms_ = false;
        bool multistate() const override { return ms_; }
        bool unity() const override { return false; }
        int states() const override { return S_; }

      public:
        SplitSum *NewInstance() const override { return new SplitSum(); }
        const string &Name() const override {
            static const string n = "OSum";
            return n;
        }
        void Set(const std::vector<const DeviceOperator *> &ops, int S) {
            ops_ = ops;
            S_ = S;
            ms_ = S > 1;
        }
        void AddTerms(qdx_ham *h, int row, int col) const override {
            for (auto *o : ops_) {
                if (auto *m = dynamic_cast<const OMultistate *>(o)) m->AddTerms(h, 0, 0);
                else o->AddTerms(h, row, col);
            }
        }
    };
    std::unique_ptr<SplitSum> sum_;

  public:
    OGSPO *NewInstance() const override { return new OGSPO(); }
    const string &Name() const override {
        static const string n = "OGSPO";
        return n;
    }
    // operators "A", "B", "C", ... in this order (OGSPO.cpp:49-61)
    void Add(const string &key, Operator *op) { ops_[key] = op; }
    void Precision(int p) { precision_ = p; }
    using Operator::Init;
    void Init(WaveFunction *psi) override {
        if (ops_.empty()) throw EParamProblem("GSPO needs at least one operator");
        check_init(psi);
        std::vector<qdx_ham *> h;
        std::vector<const DeviceOperator *> dev;
        for (auto &kv : ops_) {
            DeviceOperator *d = as_device(kv.second, "OGSPO");
            d->Clock(clock_);
            d->Init(psi);
            h.push_back(d->ham());
            dev.push_back(d);
        }
        if (prop_) qdx_prop_destroy(prop_);
        prop_ = nullptr;
        if (imaginary_) check(qdx_gspo_create_imag(h.data(), (int32_t)h.size(), clock_->Dt(), &prop_));
        else check(qdx_gspo_create(h.data(), (int32_t)h.size(), clock_->Dt(), precision_, &prop_));
        sync_direction();
        sum_.reset(new SplitSum());
        sum_->Set(dev, psi->States());
        sum_->Clock(clock_);
        sum_->Init(psi);
        H_ = sum_.get();
    }
};

}  // namespace qdxlib

#endif
