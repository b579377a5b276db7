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
//   OCheby (qdlib/OCheby.h)                      OCheby           Add("hamiltonian", H)
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
    virtual ~WaveFunction() {
        if (wf_) qdx_wf_destroy(wf_);
    }
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
    WFGridCartesian *NewInstance() const {
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
    WFMultistate *NewInstance() const {
        auto *w = new WFMultistate();
        w->alloc(g_, S_);
        return w;
    }
    // element x of state s
    dcomplex &at(int s, long x) { return (*this)[(long)s * n_ + x]; }
};

// ---------------------------------------------------------------------------
// operators
class Operator {
  protected:
    ParamContainer params_;
    QDClock *clock_ = nullptr;

  public:
    virtual ~Operator() = default;
    virtual const string &Name() const = 0;
    virtual void Init(ParamContainer &pm) { params_ = pm; }
    ParamContainer &Params() { return params_; }
    void Clock(QDClock *c) { clock_ = c; }
    QDClock *Clock() const { return clock_; }
};

// grid terms are leaves: they add themselves to a qdx Hamiltonian element
class GridOperator : public Operator {
  public:
    virtual void AddTo(qdx_ham *h, int row, int col) const = 0;
};

class OGridNablaSq : public GridOperator {
    std::vector<double> mass_;

  public:
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
    void AddTo(qdx_ham *h, int row, int col) const override {
        if (row != col) throw EIncompatible("kinetic operator on an off-diagonal element");
        std::vector<double> m(mass_);
        m.resize(4, -1.0);
        check(qdx_ham_add_kinetic(h, row, m.data()));
    }
};

class OGridPotential : public GridOperator {
  protected:
    std::vector<double> v_;

  public:
    const string &Name() const override {
        static const string n = "OGridPotential";
        return n;
    }
    void SetData(const std::vector<double> &v) { v_ = v; }
    double &operator[](long i) { return v_[i]; }
    void AddTo(qdx_ham *h, int row, int col) const override {
        if (v_.empty()) throw EParamProblem("GridPotential without data");
        check(qdx_ham_add_potential(h, row, col, v_.data()));
    }
};

// -mu(x) E(t) with E from the laser samples at the clock's step
class OGridDipole : public OGridPotential {
    std::vector<double> laser_;
    double scale_ = 1.0;

  public:
    const string &Name() const override {
        static const string n = "OGridDipole";
        return n;
    }
    std::vector<double> &GetLaser() { return laser_; }
    void Scale(double s) { scale_ = s; }
    void AddTo(qdx_ham *h, int row, int col) const override {
        if (v_.empty() || laser_.empty()) throw EParamProblem("GridDipole without dipole or laser");
        check(qdx_ham_add_dipole(h, row, col, v_.data(), laser_.data(), (int32_t)laser_.size(), scale_));
    }
};

// composite operators build a qdx Hamiltonian on first use
class OComposite : public Operator {
  protected:
    qdx_ham *ham_ = nullptr;
    qdx_grid g_{};
    int S_ = 1;
    virtual void build(qdx_ham *h) const = 0;
    virtual bool multistate() const = 0;
    virtual bool unity() const { return true; }
    virtual int states() const { return 1; }

  public:
    ~OComposite() override {
        if (ham_) qdx_ham_destroy(ham_);
    }
    // Operator::Init(WaveFunction*): bind to the wave function's grid
    void Init(WaveFunction *psi) {
        if (ham_) return;
        g_ = psi->grid();
        S_ = states();
        if (psi->States() != S_) throw EIncompatible("operator and wave function differ in the number of states");
        check(qdx_ham_create(context(), &g_, S_, multistate() ? 1 : 0, unity() ? 1 : 0, &ham_));
        build(ham_);
    }
    using Operator::Init;
    qdx_ham *ham() const { return ham_; }
    dcomplex Emax() {
        double lo, hi;
        check(qdx_ham_bounds(ham_, &lo, &hi));
        return hi;
    }
    dcomplex Emin() {
        double lo, hi;
        check(qdx_ham_bounds(ham_, &lo, &hi));
        return lo;
    }
    void Apply(WaveFunction *dst, WaveFunction *src) {
        Init(src);
        check(qdx_ham_apply(ham_, dst->device(), src->device(), clock_ ? clock_->TimeStep() : 0));
        dst->DeviceChanged();
    }
    // <psi|H|psi>
    dcomplex Expec(WaveFunction *psi) {
        Init(psi);
        double e = 0;
        check(qdx_ham_expec(ham_, psi->device(), clock_ ? clock_->TimeStep() : 0, &e));
        return e;
    }
};

class OSum : public OComposite {
    std::vector<const GridOperator *> ops_;
    void build(qdx_ham *h) const override {
        if (ops_.empty()) throw EParamProblem("Sum Operator is empty");
        for (auto *o : ops_) o->AddTo(h, 0, 0);
    }
    bool multistate() const override { return false; }

  public:
    const string &Name() const override {
        static const string n = "OSum";
        return n;
    }
    void Add(const GridOperator *o) { ops_.push_back(o); }
    int Size() const { return (int)ops_.size(); }
};

class OMultistate : public OComposite {
    struct E {
        const GridOperator *op;
        int i, j;
    };
    std::vector<E> elems_;
    void build(qdx_ham *h) const override {
        if (elems_.empty()) throw EParamProblem("Multistate operator is empty");
        for (auto &e : elems_) e.op->AddTo(h, e.i, e.j);
    }
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
    const string &Name() const override {
        static const string n = "OMultistate";
        return n;
    }
    void Add(const GridOperator *o, int i, int j) { elems_.push_back({o, i, j}); }
};

// ---------------------------------------------------------------------------
// propagators (OPropagator, OCheby, OGSPO)
class OPropagator : public Operator {
  protected:
    qdx_prop *prop_ = nullptr;

  public:
    ~OPropagator() override {
        if (prop_) qdx_prop_destroy(prop_);
    }
    // one time step at the clock's step
    void Apply(WaveFunction *psi) {
        if (!prop_) Init(psi);
        qdx_step_stats st;
        check(qdx_prop_apply(prop_, psi->device(), clock_ ? clock_->TimeStep() : 0, &st));
        psi->DeviceChanged();
    }
    void Apply(WaveFunction *dst, WaveFunction *src) {
        *dst = *src;
        Apply(dst);
    }
    virtual void Init(WaveFunction *psi) = 0;
    using Operator::Init;
};

class OCheby : public OPropagator {
    OComposite *H_ = nullptr;
    int precision_ = QDX_FP64;

  public:
    const string &Name() const override {
        static const string n = "OCheby";
        return n;
    }
    // Add("hamiltonian", H)
    void Add(const string &key, OComposite *H) {
        if (key != "hamiltonian") throw EParamProblem("OCheby takes a hamiltonian");
        H_ = H;
    }
    void Precision(int p) { precision_ = p; }
    using Operator::Init;
    // params "order", "prec", "scaling" as OCheby::Init(ParamContainer&)
    void Init(WaveFunction *psi) override {
        if (!H_) throw EParamProblem("Chebychev propagator has no hamiltonian");
        if (!clock_) throw EParamProblem("No clock");
        H_->Init(psi);
        int order = 0;
        double prec = -1, scaling = 0;
        params_.GetValue("order", order);
        params_.GetValue("prec", prec);
        params_.GetValue("scaling", scaling);
        if (prop_) qdx_prop_destroy(prop_);
        prop_ = nullptr;
        check(qdx_cheby_create(H_->ham(), clock_->Dt(), order, prec, scaling, precision_, &prop_));
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

class OGSPO : public OPropagator {
    std::map<string, OComposite *> ops_;
    int precision_ = QDX_FP64;

  public:
    const string &Name() const override {
        static const string n = "OGSPO";
        return n;
    }
    // operators "A", "B", "C", ... in this order (OGSPO.cpp:49-61)
    void Add(const string &key, OComposite *op) { ops_[key] = op; }
    void Precision(int p) { precision_ = p; }
    using Operator::Init;
    void Init(WaveFunction *psi) override {
        if (ops_.empty()) throw EParamProblem("GSPO needs at least one operator");
        if (!clock_) throw EParamProblem("No clock");
        std::vector<qdx_ham *> h;
        for (auto &kv : ops_) {
            kv.second->Init(psi);
            h.push_back(kv.second->ham());
        }
        if (prop_) qdx_prop_destroy(prop_);
        prop_ = nullptr;
        check(qdx_gspo_create(h.data(), (int32_t)h.size(), clock_->Dt(), precision_, &prop_));
    }
};

}  // namespace qdxlib

#endif
