"""Program inputs (QDng's XML `.qdi` format) for the end-to-end tests of
qdng_amd.qdi, generated from parameter tables.

Each table row names one regression case of QDng's test suite (the names
tests/checktests.sh greps for) and holds the physical parameters of that case:
time step, step count, write cycle, propagator and its options, the operators
of the Hamiltonian with their masses, data files and keys, the initial wave
function, and the filters whose values the checks read.  The XML is built
here element by element (`@srcdir@` stands for the folder of the data files,
as in QDng's inputs), so the tests need no copy of QDng's input programs.

    python tests/golden/qdi_programs.py OUTDIR   # writes OUTDIR/<case>.qdi.in
"""
import os
import sys
import xml.etree.ElementTree as ET

MASS = "1763.2"          # the reduced mass of every grid case
SRC = "@srcdir@/"

# ---------------------------------------------------------------------------
# parameter tables

# real-time propagations: (dt, steps, wcycle), propagator (name, options),
# Hamiltonian kind, initial wave function, optional init filter, post filters
PROPAGATIONS = {
    "TGP-Expec": dict(dt=20, steps=1000, wcycle=1000, prop=("Cheby", {"prec": "1e-12"}), ham="1state",
                      wf="ef0", post=("1state", [{"num": "4"}])),
    "TGP-Kick": dict(dt=20, steps=1000, wcycle=1000, prop=("Cheby", {"prec": "1e-12"}), ham="1state",
                     wf="ef0", init=("GridKick", {"k": "1"}), post=("1state", [{"num": "4"}])),
    "TGPL-IR-Pulse-Cheby": dict(dt=10, steps=6750, wcycle=6750, prop=("Cheby", {"prec": "1e-12"}),
                                ham="1state+dipole", laser="GridLaser_w0.00713", wf="ef0",
                                post=("1state", [{"start": "0"}, {"start": "1"}, {"start": "2"}])),
    "TGPL-IR-Pulse-GSPO": dict(dt=10, steps=6750, wcycle=6750, prop=("GSPO", {}), ham="1state+dipole",
                               laser="GridLaser_w0.00713", wf="ef0",
                               post=("1state", [{"start": "0"}, {"start": "1"}, {"start": "2"}])),
    "TGPMS-Cheby": dict(dt=10, steps=1000, wcycle=1000, prop=("Cheby", {"prec": "1e-12"}), ham="2state",
                        wf="ef0 on both states", post=("2state", None)),
    "TGPMS-GSPO": dict(dt=2, steps=3000, wcycle=3000, prop=("GSPO", {}), ham="2state diagonal",
                       wf="ef0 on both states", post=("2state", None)),
    "TGPLMS-Cheby": dict(dt=1, steps=400, wcycle=1000, prop=("Cheby", {"prec": "1e-12"}), ham="2state+dipole",
                         laser="GridLaser_w0.176", wf="ef0 on state 0 of 2", post=("2state", None)),
    "TGPLMS-SIL": dict(dt=1, steps=400, wcycle=1000, prop=("SIL", {"order": "20"}), ham="2state+dipole",
                       laser="GridLaser_w0.176", wf="ef0 on state 0 of 2", post=("2state", None)),
    # absorbing boundary: T + a negative imaginary potential (Gobbler, high
    # pass at x = 5.5) + a real Gobbler wall (low pass at x = 2)
    "TGP-NIP-Cheby": dict(dt=10, steps=1000, wcycle=1000, prop=("Cheby", {}), ham="1state+nip", wf="ef0",
                          post=("1state, no projections", None)),
}
# the Gobbler operators of 1state+nip: (tag, options)
NIP_TERMS = [("Vpot", {"dims": "1", "rp0": "5.5", "order": "5", "gain": "0.02", "nip": "true"}),
             ("VpotIm", {"dims": "1", "lp0": "2", "order": "20", "gain": "0.3", "nip": "false", "key": "V"})]

# eigenstate programs: an imaginary-time relaxation of 10 states from the
# Gaussian, then the autocorrelation method on the sum of the first six
# states the relaxation wrote
EIGEN = {
    "TGE-Cheby": ("Cheby", {}),
    "TGE-GSPO": ("GSPO", {}),
    "TGE-SIL": ("SIL", {"order": "25"}),
    "TGE-Arnoldi": ("Arnoldi", {"order": "15"}),
}
# optimal control programs: OCT attributes, Hamiltonian kind and laser, the
# initial state, the target (overlap with an eigenfunction, or a projection
# onto eigenfunctions: (files, start, num), on state 1 for the two-state
# case); the one-state cases follow with a GSPO propagation under the
# optimised laser, as QDng's inputs do
OCT = {
    "TGO-IR-Krotov": dict(attrs=dict(alpha=400, dt=10, steps=6750, shape="GridShapeIR", method="krotov"),
                          ham="1state+dipole", laser="GridLaser_w0.00713", target=("wft", "GridPot0_ef_1")),
    "TGO-IR-Rabitz": dict(attrs=dict(alpha=50, dt=10, steps=6750, shape="GridShapeIR", iterations=8,
                                     method="rabitz"),
                          ham="1state+dipole", laser="GridLaser_w0.00713", target=("wft", "GridPot0_ef_1")),
    "TGO-IR-Krotov-Projection": dict(attrs=dict(alpha=400, dt=10, steps=6750, shape="GridShapeIR",
                                                method="krotov", ttype="operator"),
                                     ham="1state+dipole", laser="GridLaser_w0.00713",
                                     target=("projection", ("GridPot0_ef", 1, 1))),
    "TGO-UV-Krotov": dict(attrs=dict(alpha=10, dt=1, steps=400, shape="GridShapeUV", method="krotov",
                                     ttype="operator", iterations=20),
                          ham="2state+dipole", laser="GridLaser_w0.176",
                          target=("projection on state 1", ("GridPot0_ef", 0, 4))),
}
OCT_PROPA = dict(dt=10, steps=6750, wcycle=6750)

EIGEN_IMAG = dict(dt=5, steps=1000, Nef=10)
EIGEN_AC = dict(dt=5, steps=2048, lc=6)


# ---------------------------------------------------------------------------
# element builders

def _el(tag, parent=None, **attrs):
    e = ET.Element(tag) if parent is None else ET.SubElement(parent, tag)
    for k, v in attrs.items():
        e.set(k, str(v))
    return e


def _kinetic(parent, tag, key=None):
    e = _el(tag, parent, name="GridNablaSq", mass=MASS)
    if key:
        e.set("key", key)
    return e


def _potential(parent, tag, fname, key=None):
    e = _el(tag, parent, name="GridPotential", file=SRC + fname)
    if key:
        e.set("key", key)
    return e


def _dipole(parent, tag, fname, laser):
    return _el(tag, parent, name="GridDipole", file=SRC + fname, laser=SRC + laser)


def _one_state_terms(parent, keys, laser):
    _kinetic(parent, "A", "T" if keys else None)
    _potential(parent, "B", "GridPot0", "V" if keys else None)
    if laser:
        _dipole(parent, "C", "GridDip0", laser)


def _propagator(parent, spec, laser):
    name, opts = spec["prop"]
    p = _el("propagator", parent, name=name, **opts)
    ham = spec["ham"]
    if ham == "1state+nip":
        h = _el("hamiltonian", p, name="Sum")
        _kinetic(h, "A", "T")
        for tag, opts in NIP_TERMS:
            _el(tag, h, name="Gobbler", **opts)
    elif ham.startswith("1state"):
        # GSPO takes the split operators directly, the others one Hamiltonian
        target = p if name == "GSPO" else _el("hamiltonian", p, name="Sum")
        _one_state_terms(target, True, laser)
    elif ham == "2state diagonal":
        a = _el("A", p, name="DMultistate")
        _kinetic(a, "m0.0", "T")
        _kinetic(a, "m1.1", "T1")
        b = _el("B", p, name="DMultistate")
        _potential(b, "m0.0", "GridPot0", "V0")
        _potential(b, "m1.1", "GridPot1", "V1")
    else:
        h = _el("hamiltonian", p, name="Multistate")
        for s in (0, 1):
            d = _el("m%d.%d" % (s, s), h, name="Sum")
            if s == 0:
                _kinetic(d, "Tkin", "T")
            else:
                _el("Tkin", d, ref="T")
            _potential(d, "Vpot", "GridPot%d" % s, "V%d" % s)
        if laser:
            _dipole(h, "m1.0", "GridDip01", laser)
    return p


def _wavefunction(parent, kind):
    if kind == "ef0":
        return _el("wf", parent, file=SRC + "GridPot0_ef_0", normalize="true")
    if kind == "ef0 on both states":
        w = _el("wf", parent, name="Multistate", normalize="true")
        for s in (0, 1):
            _el("wf%d" % s, w, file=SRC + "GridPot0_ef_0")
        return w
    w = _el("wf", parent, name="Multistate", normalize="true", states="2")
    _el("wf0", w, file=SRC + "GridPot0_ef_0")
    return w


def _post_filters(parent, kind, projections):
    f = _el("filterpost", parent)
    if kind == "1state, no projections":
        for cls in ("GridMomentum", "GridPosition"):
            _el("expeconly", f, name=cls)
        _el("expeconly", f, ref="T")
        _el("expeconly", f, ref="V")
        return f
    if kind == "1state":
        _el("expeconly", f, name="GridMomentum")
        _el("expeconly", f, name="GridPosition")
        for pr in projections:
            _el("expeconly", f, name="Projection", files=SRC + "GridPot0_ef", **pr)
        _el("expeconly", f, ref="T")
        _el("expeconly", f, ref="V")
        return f
    # per-state expectation values through two-state diagonal wrappers:
    # momentum, position, kinetic and potential energy of states 0 and 1
    for first, refs in (("GridMomentum", ("p", "p")), ("GridPosition", ("x", "x")), (None, ("T", "T")),
                        (None, ("V0", "V1"))):
        for s in (0, 1):
            m = _el("expeconly", f, name="Multistate", unity="false", states="2")
            if s == 0 and first:
                _el("m0.0", m, name=first, key=refs[0])
            else:
                _el("m%d.%d" % (s, s), m, ref=refs[s])
    return f


def propagation(name):
    spec = PROPAGATIONS[name]
    root = _el("qdng")
    pr = _el("propa", root, dt=spec["dt"], steps=spec["steps"], wcycle=spec["wcycle"], dir=name,
             spectrum="spectrum.dat")
    _propagator(pr, spec, spec.get("laser"))
    _wavefunction(pr, spec["wf"])
    if "init" in spec:
        fi = _el("filterinit", pr)
        cls, opts = spec["init"]
        _el("apply", fi, name=cls, **opts)
    kind, projections = spec["post"]
    _post_filters(pr, kind, projections)
    return root


def eigen(name):
    pname, opts = EIGEN[name]
    root = _el("qdng")
    for method, par in (("imag", EIGEN_IMAG), ("ac", EIGEN_AC)):
        attrs = dict(dt=par["dt"], steps=par["steps"])
        if method == "imag":
            attrs["Nef"] = par["Nef"]
        e = _el("eigen", root, dir="%s-%s" % (name, "im" if method == "imag" else "ac"), method=method, **attrs)
        p = _el("propagator", e, name=pname, **opts)
        _one_state_terms(p if pname == "GSPO" else _el("hamiltonian", p, name="Sum"), False, None)
        if method == "imag":
            _el("wf", e, file=SRC + "GridGauss", normalize="true")
        else:
            w = _el("wf", e, name="LC", normalize="true")
            for k in range(par["lc"]):
                _el("wf%d" % k, w, file="%s-im/ef_%d" % (name, k))
    return root


def oct_program(name):
    spec = OCT[name]
    root = _el("qdng")
    attrs = dict(spec["attrs"])
    attrs["shape"] = SRC + attrs["shape"]
    o = _el("oct", root, dir=name, **attrs)
    _propagator(o, dict(prop=("Cheby", {"prec": "1e-12"}), ham=spec["ham"]), spec["laser"])
    if spec["ham"].startswith("1state"):
        _el("wfi0", o, name="GridCartesian", file=SRC + "GridPot0_ef_0", normalize="true")
    else:
        w = _el("wfi0", o, name="Multistate", normalize="true", states="2")
        _el("wf0", w, file=SRC + "GridPot0_ef_0")
    kind, what = spec["target"]
    if kind == "wft":
        _el("wft0", o, name="GridCartesian", file=SRC + what, normalize="true")
    elif kind == "projection":
        files, start, num = what
        _el("target0", o, name="Projection", files=SRC + files, start=start, num=num)
    else:
        files, start, num = what
        t = _el("target0", o, name="Multistate", unity="false", states="2")
        _el("m1.1", t, name="Projection", files=SRC + files, start=start, num=num)
    if spec["ham"].startswith("1state"):
        pr = _el("propa", root, dir=name, nfile="norm", spectrum="spectrum.dat", **OCT_PROPA)
        p = _el("propagator", pr, name="GSPO")
        _kinetic(p, "A", "T")
        _potential(p, "B", "GridPot0", "V")
        _el("C", p, name="GridDipole", file=SRC + "GridDip0", laser=name + "/laser")
        _wavefunction(pr, "ef0")
        _post_filters(pr, "1state", [{"start": "0"}, {"start": "1"}, {"start": "2"}])
    return root


def text(name):
    if name in PROPAGATIONS:
        root = propagation(name)
    elif name in OCT:
        root = oct_program(name)
    else:
        root = eigen(name)
    ET.indent(root, "  ")
    return ET.tostring(root, encoding="unicode") + "\n"


def write(name, outdir):
    path = os.path.join(outdir, name + ".qdi.in")
    with open(path, "w") as f:
        f.write(text(name))
    return path


def write_all(outdir):
    os.makedirs(outdir, exist_ok=True)
    return {n: write(n, outdir) for n in list(PROPAGATIONS) + list(EIGEN) + list(OCT)}


if __name__ == "__main__":
    for n, p in write_all(sys.argv[1] if len(sys.argv) > 1 else ".").items():
        print(p)
