"""QDng program input files (XML, `.qdi`) run on the device path.

The loader follows QDng's operator-graph construction (src/ChainLoader.cpp:
34-457): an element's `name` attribute names the class, its tag is free
(`A`, `Tkin`, `m1.0`, ...), `key` registers the object and `ref` reuses it;
`Sum` adds its children, `Multistate`/`DMultistate` place theirs at the matrix
element named by the tag `m<i>.<j>`.  Supported classes: GridNablaSq
(`mass`, comma separated per dimension), GridPotential (`file`, `setzero`,
`offset`, `scale`; OGridPotential.cpp:36-76), GridDipole (`file`, `laser`,
`scale`), Gobbler (`lp<d>`, `rp<d>`, `order` (20), `gain` (1), `nip`, `dims`;
OGobbler.cpp:85-164), Sum, Multistate (`unity`, `states`), DMultistate; as filters also
GridMomentum and GridPosition (`dim`) and Projection (`files`, `start`, `num`,
`step`; OProjection.cpp:42-68).  Propagators: Cheby (`prec`, `order`,
`scaling`), GSPO (split operators as children), SIL (`order`, `err`), Arnoldi
(`order`, `conv`).  Wave functions: a single `file`, or `name="Multistate"`
with children `wf<i>` (missing states are zero), `normalize`.

Programs:
  propa  ProgPropa (src/ProgPropa.cpp:40-160, 270-340): `dt`, `steps`,
         `wcycle` (10), `dir`, `fname` ("WF"), `nfile`; the report table of
         Reporter::Analyze (src/Reporter.cpp:63-158, taken before each step:
         step, time, population, populations per state, energy, <psi0|psi>),
         pre- and post-step filters written to `expec_pre.dat` and
         `expec_post.dat`
         (src/FiltersList.cpp:135-330, values renormalised by <psi|psi>), the
         wave function written every `wcycle` steps with Propagation.meta.
  eigen  ProgEigen (src/ProgEigen.cpp:50-80, 150-260) method "imag" (`Nef`
         (20), `conv`, `diag` on) or "ac" (`win` on, `tol` 1, `width`, `diag`
         off; src/ProgEigen.cpp:414-600): `dt`, `steps` (512), `fname`
         ("ef"), `ename` ("energy.dat"), `dir`; eigenfunctions written as a
         sequence and the energy table in QDng's format.  Relative wave
         function files are looked up in the output folder first (a later
         program reads what an earlier one wrote).

  oct    ProgOCT (src/ProgOCT.cpp:59-300, 868-1092): `dt`, `steps`, `shape`
         (laser file of the shape function), `method` (krotov, rabitz,
         rabitzfb), `ttype` (overlap: `wft<n>` target functions; operator:
         `target<n>` Projections, or a Multistate of diagonal Projections
         with unity off), `phase`, `ntargets`, `alpha`, `iterations` (50),
         `conv` (1e-5), `fname` ("laser"), `writeyield`/`yfile`; `wfi<n>`
         initial functions, an optional `gobbler` operator; the optimised
         laser is written as `dir/fname` and the iteration table of
         ProgOCT::Report is logged (or written to the yield file).  Not
         supported: method freq, moving targets (mv), writel.

With `spectrum=` both programs write the autocorrelation spectrum
(Reporter::WriteSpectrum, src/Reporter.cpp:160-190; Hann window unless
`win="false"`).  `LC` wave functions sum their members as ChainLoader::LoadWFLC_ does (each
twice, no coefficient).  Not supported (QdxError with QDX_ERR_INCOMPAT): other
operator classes, reading a stored propagation (`read=`), filter actions
other than `expeconly`, `normalize` and `apply` of a GridKick.  A
propagation `start` = k restarts as ProgPropa does: the clock (laser sample
and report time) begins at step k, the loop index runs k .. steps (so
steps − k + 1 steps), the loaded state is not written again and written
files are numbered from k.  Init filters (`filterinit`) run once
before the initial state is kept for the autocorrelation.
"""
from __future__ import annotations

import argparse
import math
import os
import re
import sys
import xml.etree.ElementTree as ET

import numpy as np

from . import (Arnoldi, Cheby, GSPO, Hamiltonian, Projection, QdxError, SIL, WaveFunction, eigen_ac, eigen_imag,
               get_file_format, load_laser, oct_run, save_laser, set_file_format, write_propagation_meta)

ERR_PARAM, ERR_INCOMPAT = -1, -2
AU2WAVENUMBERS = 219475.0   # src/qdlib/Conversion.h:9


def _err(msg, code=ERR_PARAM):
    return QdxError(code, msg)


def _bool(v, default=False):
    if v is None:
        return default
    s = v.strip().lower()
    if s in ("true", "1", "yes", "on"):
        return True
    if s in ("false", "0", "no", "off"):
        return False
    raise _err(f"not a boolean: '{v}'")


def _attrs(el, srcdir):
    a = dict(el.attrib)
    if srcdir is not None:
        a = {k: v.replace("@srcdir@", srcdir) for k, v in a.items()}
    return a


# ---------------------------------------------------------------------------
# operator graph

class Term:
    """One grid operator: kind 'kin', 'pot', 'dip' (Hamiltonian terms) or
    'mom', 'pos', 'proj' (filters only)."""

    def __init__(self, kind, cls, **params):
        self.kind, self.cls, self.params = kind, cls, params


class Sum:
    def __init__(self, items):
        self.items, self.cls = items, "Sum"


class Multi:
    def __init__(self, elems, states, unity, dmulti):
        self.elems, self.states, self.unity, self.dmulti = elems, states, unity, dmulti
        self.cls = "DMultistate" if dmulti else "Multistate"

    @property
    def S(self):
        n = max(max(i, j) for i, j in self.elems) + 1 if self.elems else 1
        return max(n, self.states)


class Loader:
    """ChainLoader: operators from XML elements with key/ref bookkeeping."""

    def __init__(self, srcdir=None, base=None):
        self.srcdir = srcdir
        self.base = base      # the output folder: relative data files are looked up there first
        self.keys = {}

    def op(self, el):
        a = _attrs(el, self.srcdir)
        if "ref" in a:
            if a["ref"] not in self.keys:
                raise _err(f"operator reference '{a['ref']}' is not defined")
            return self.keys[a["ref"]]
        name = a.get("name")
        if not name:
            raise _err(f"element <{el.tag}> names no operator class")
        if name == "Sum":
            o = Sum([self.op(c) for c in el])
            if not o.items:
                raise _err("empty Sum")
        elif name in ("Multistate", "DMultistate"):
            elems = {}
            for c in el:
                m = re.fullmatch(r"m(\d+)\.(\d+)", c.tag)
                if not m:
                    raise _err(f"matrix element tag <{c.tag}> is not m<i>.<j>")
                elems[(int(m.group(1)), int(m.group(2)))] = self.op(c)
            o = Multi(elems, int(a.get("states", 0)), _bool(a.get("unity"), True), name == "DMultistate")
        elif name == "GridNablaSq":
            if "mass" not in a:
                raise _err("GridNablaSq needs a mass")
            o = Term("kin", name, mass=[float(v) for v in a["mass"].split(",")])
        elif name == "GridPotential":
            o = Term("pot", name, file=self._file(a, name), setzero=_bool(a.get("setzero")),
                     offset=float(a.get("offset", 0.0)), scale=float(a.get("scale", 1.0)))
        elif name == "GridDipole":
            if "laser" not in a:
                raise _err("GridDipole needs a laser file")
            o = Term("dip", name, file=self._file(a, name), laser=_resolve(a["laser"], self.base),
                     scale=float(a.get("scale", 1.0)))
        elif name == "Gobbler":
            order = int(a.get("order", 20))
            if order < 1:
                raise _err("Filter order to small")
            if _bool(a.get("residue")):
                raise _err("Gobbler: residue expectation values are not supported", ERR_INCOMPAT)
            edges = {}
            for side in ("lp", "rp"):
                edges[side] = {int(k[2:]): float(v) for k, v in a.items() if re.fullmatch(side + r"\d+", k)}
            o = Term("gob", name, lp=edges["lp"], rp=edges["rp"], order=order, gain=float(a.get("gain", 1.0)),
                     nip=_bool(a.get("nip")), dims=int(a["dims"]) if "dims" in a else None)
        elif name in ("GridMomentum", "GridPosition"):
            o = Term("mom" if name == "GridMomentum" else "pos", name, dim=int(a.get("dim", 0)))
        elif name == "GridKick":
            if "k" not in a:
                raise _err("GridKick needs a momentum vector k")
            o = Term("kick", name, k=[float(v) for v in a["k"].split(",")], scale=float(a.get("scale", 1.0)))
        elif name == "Projection":
            o = Term("proj", name, files=self._file(a, name, "files"), start=int(a.get("start", 0)),
                     num=int(a.get("num", 1)), step=int(a.get("step", 1)))
        else:
            raise _err(f"operator class '{name}' is not supported", ERR_INCOMPAT)
        if "key" in a:
            self.keys[a["key"]] = o
        return o

    def _file(self, a, name, attr="file"):
        if attr not in a:
            raise _err(f"{name} needs a '{attr}' attribute")
        return _resolve(a[attr], self.base)

    def propagator(self, el):
        a = _attrs(el, self.srcdir)
        name = a.get("name")
        if name in ("Cheby", "SIL", "Arnoldi"):
            hs = [c for c in el if c.tag == "hamiltonian"]
            if len(hs) != 1:
                raise _err(f"{name} needs one <hamiltonian>")
            return name, a, [self.op(hs[0])]
        if name == "GSPO":
            ops = [self.op(c) for c in el]
            if len(ops) < 2:
                raise _err("GSPO needs at least two split operators")
            return name, a, ops
        raise _err(f"propagator '{name}' is not supported", ERR_INCOMPAT)


def _terms(op):
    """Flatten a Sum (or a single term) into Hamiltonian terms."""
    if isinstance(op, Term):
        if op.kind not in ("kin", "pot", "dip", "gob"):
            raise _err(f"{op.cls} is a filter, not a Hamiltonian term", ERR_INCOMPAT)
        return [op]
    if isinstance(op, Sum):
        return [t for o in op.items for t in _terms(o)]
    raise _err(f"{op.cls} inside a matrix element is not supported", ERR_INCOMPAT)


def _add_term(H, i, j, t):
    p = t.params
    if t.kind == "kin":
        if i != j:
            raise _err("a kinetic operator belongs on the diagonal", ERR_INCOMPAT)
        H.add_kinetic(i, p["mass"])
    elif t.kind == "pot":
        H.add_potential_file(i, j, p["file"], p["setzero"], p["offset"], p["scale"])
    elif t.kind == "gob":
        nd = len(H.N)
        if p["dims"] is not None and p["dims"] != nd:
            raise _err("OGobbler: Number of dims doesn't match WF", ERR_INCOMPAT)
        # lp<d> / rp<d> beyond the wave function's dimensions are not read
        H.add_gobbler(i, j, {d: x for d, x in p["lp"].items() if d < nd},
                      {d: x for d, x in p["rp"].items() if d < nd}, p["order"], p["gain"], p["nip"])
    else:
        H.add_dipole_file(i, j, p["file"], p["laser"], p["scale"])


def _grid_lists(w):
    g = w.grid
    return list(w.N), [g.xmin[d] for d in range(len(w.N))], [g.xmax[d] for d in range(len(w.N))]


def build_hamiltonian(ctx, op, psi):
    """A device Hamiltonian (OSum or OMultistate) on the grid of psi."""
    N, lo, hi = _grid_lists(psi)
    if isinstance(op, Multi):
        if op.S != psi.S:
            raise _err(f"{op.cls} has {op.S} states, the wave function {psi.S}", ERR_INCOMPAT)
        H = Hamiltonian(ctx, N, lo, hi, op.S, True, op.unity)
        for (i, j), e in sorted(op.elems.items()):
            for t in _terms(e):
                _add_term(H, i, j, t)
        return H
    if psi.S != 1:
        raise _err("a multistate wave function needs a Multistate operator", ERR_INCOMPAT)
    H = Hamiltonian(ctx, N, lo, hi)
    for t in _terms(op):
        _add_term(H, 0, 0, t)
    return H


def _merged(ops):
    """The Hamiltonian of a GSPO propagator: the sum of its split operators."""
    if all(isinstance(o, Multi) for o in ops):
        elems = {}
        for o in ops:
            for k, e in o.elems.items():
                elems.setdefault(k, []).append(e)
        return Multi({k: Sum(v) for k, v in elems.items()}, max(o.S for o in ops), True, False)
    if any(isinstance(o, Multi) for o in ops):
        raise _err("GSPO split operators mix Multistate and single-state operators", ERR_INCOMPAT)
    return Sum(list(ops))


def build_propagator(ctx, spec, psi, dt, imag=False):
    """(propagator, Hamiltonian of the energies) from Loader.propagator()."""
    name, a, ops = spec
    if name == "GSPO":
        U = GSPO([build_hamiltonian(ctx, o, psi) for o in ops], dt, imag=imag)
        return U, build_hamiltonian(ctx, _merged(ops), psi)
    H = build_hamiltonian(ctx, ops[0], psi)
    if name == "Cheby":
        U = Cheby(H, dt, order=int(a.get("order", 0)), prec=float(a.get("prec", -1.0)),
                  scaling=float(a.get("scaling", 0.0)), imag=imag)
    elif name == "SIL":
        U = SIL(H, dt, order=int(a.get("order", 0)), err=float(a.get("err", 0.0)), imag=imag)
    else:
        U = Arnoldi(H, dt, order=int(a.get("order", 0)), conv=float(a.get("conv", 0.0)), imag=imag)
    return U, H


def _resolve(path, base):
    """A relative file name is looked up under `base` (the output folder)
    first: programs read what earlier programs of the same input wrote."""
    if not base or os.path.isabs(path):
        return path
    cand = os.path.join(base, path)
    metas = [cand + ".meta", cand + ".wf", cand + ".op"]
    if "_" in os.path.basename(cand):
        metas.append(cand.rsplit("_", 1)[0] + ".meta")
    return cand if any(os.path.exists(m) for m in metas) else path


def load_wave_function(ctx, el, srcdir=None, base=None):
    a = _attrs(el, srcdir)
    name = a.get("name")
    if name in (None, "GridCartesian", "WFGridCartesian"):   # ModuleLoader adds the "WF" prefix
        if "file" not in a:
            raise _err("<wf> needs a file")
        w = WaveFunction.load(ctx, _resolve(a["file"], base))
    elif name == "Multistate":
        parts = {}
        for c in el:
            m = re.fullmatch(r"wf(\d+)", c.tag)
            if not m:
                raise _err(f"state tag <{c.tag}> is not wf<i>")
            parts[int(m.group(1))] = load_wave_function(ctx, c, srcdir, base)
        if not parts:
            raise _err("empty Multistate wave function")
        S = max(int(a.get("states", 0)), max(parts) + 1)
        first = next(iter(parts.values()))
        N, lo, hi = _grid_lists(first)
        w = WaveFunction(ctx, N, lo, hi, S)
        w.upload(np.zeros(S * w.size, dtype=np.complex128))
        for s, p in parts.items():
            if p.S != 1 or _grid_lists(p) != (N, lo, hi):
                raise _err("the states of a Multistate wave function need one grid", ERR_INCOMPAT)
            w.upload(p.download(), s)
    elif name == "LC":
        # ChainLoader::LoadWFLC_ (src/ChainLoader.cpp:257-283) adds every
        # member twice and applies no coefficient
        w = None
        for c in el:
            x = load_wave_function(ctx, c, srcdir, base)
            if w is None:
                N, lo, hi = _grid_lists(x)
                w = WaveFunction(ctx, N, lo, hi, x.S)
                w.copy_from(x)
            else:
                w.axpy(1.0, x)
            w.axpy(1.0, x)
        if w is None:
            raise _err("empty LC wave function")
    else:
        raise _err(f"wave function type '{name}' is not supported", ERR_INCOMPAT)
    if _bool(a.get("normalize")):
        w.normalize()
    return w


# ---------------------------------------------------------------------------
# post-step filters (expeconly)

class _Filter:
    def __init__(self, ctx, op, psi, label):
        self.label = label
        self.kind = None
        N, lo, hi = _grid_lists(psi)
        if isinstance(op, Multi):
            # expectation of an S-state operator with (mostly) one element
            kinds = [e.kind for e in op.elems.values() if isinstance(e, Term)]
            if len(op.elems) == 1 and kinds and kinds[0] in ("mom", "pos"):
                (s, s2), e = next(iter(op.elems.items()))
                if s != s2:
                    raise _err("position/momentum filters sit on the diagonal", ERR_INCOMPAT)
                self.kind, self.state, self.dim = e.kind, s, e.params["dim"]
            else:
                if op.S != psi.S:
                    raise _err(f"filter has {op.S} states, the wave function {psi.S}", ERR_INCOMPAT)
                self.kind, self.H = "ham", build_hamiltonian(ctx, op, psi)
        elif isinstance(op, Term) and op.kind in ("mom", "pos"):
            if psi.S != 1:
                raise _err("position/momentum of a multistate function: use a Multistate filter", ERR_INCOMPAT)
            self.kind, self.state, self.dim = op.kind, 0, op.params["dim"]
        elif isinstance(op, Term) and op.kind == "proj":
            if psi.S != 1:
                raise _err("Projection filters on one state only", ERR_INCOMPAT)
            p = op.params
            self.kind = "proj"
            self.efs = [WaveFunction.load(ctx, p["files"], counter=p["start"] + k * p["step"])
                        for k in range(p["num"])]
        else:
            self.kind, self.H = "ham", build_hamiltonian(ctx, op, psi)

    def value(self, psi, step):
        if self.kind == "mom":
            return float(psi.expec_momentum(self.dim)[self.state])
        if self.kind == "pos":
            return float(psi.expec_position(self.dim)[self.state])
        if self.kind == "proj":
            return float(sum(abs(psi.overlap(e)) ** 2 for e in self.efs))
        return self.H.expec(psi, step)


def _g10(v):
    return "%.10g" % v


class _Kick:
    """OGridKick (src/qdlib/OGridKick.cpp): psi(x) *= exp(i sum_n k_n scale
    x_n), x_n = xmin_n + i dx_n; applied on the host (an initialisation
    filter, not a per-step operator)."""

    def __init__(self, op, psi):
        N, lo, hi = _grid_lists(psi)
        k = op.params["k"]
        if len(k) > len(N):
            raise _err("OGridKick: Number of dims doesn't match WF")
        if psi.S != 1:
            raise _err("GridKick acts on one state", ERR_INCOMPAT)
        ph = np.zeros(N[::-1])
        for n in range(len(N)):
            kn = (k[n] if n < len(k) else 0.0) * op.params["scale"]
            if kn != 0:
                dx = (hi[n] - lo[n]) / (N[n] - 1)
                x = (lo[n] + np.arange(N[n]) * dx) * kn
                shape = [1] * len(N)
                shape[len(N) - 1 - n] = N[n]   # dimension 0 fastest
                ph = ph + x.reshape(shape)
        self.f = np.exp(1j * ph).ravel()

    def apply(self, psi):
        psi.upload(psi.download() * self.f)


class FilterList:
    """FiltersList (src/FiltersList.cpp:135-330) for one <filterpre> or
    <filterpost> section: actions `expeconly` (write Re<psi|O|psi>/<psi|psi>)
    and `normalize`, in order; one line per call into the section's file
    (`fname`, default expec_pre.dat / expec_post.dat): the time, then the
    values, tab-terminated, %.10g.  Values are divided by <psi|psi> unless
    the section sets renorm="false".  Init sections (`filterinit`) start
    unrenormalised: QDng's FiltersList defaults to false
    (src/FiltersList.cpp:11) and only the pre and post sections switch it on
    (src/ProgPropa.cpp:260, 276); a `renorm` attribute overrides either."""

    def __init__(self, ctx, ld, sections, psi, d, default_name, srcdir, renorm_default=True):
        self.items = []
        self.f = None
        if not sections:
            return
        if len(sections) > 1:
            raise _err("one filter section of each kind")
        sa = _attrs(sections[0], srcdir)
        # the section's renorm attribute overrides the default (FiltersList.cpp:181-182)
        self.renorm = _bool(sa.get("renorm"), renorm_default)
        for f in sections[0]:
            if f.tag == "normalize":
                self.items.append(None)
                continue
            if f.tag == "apply":
                op = ld.op(f)
                if not (isinstance(op, Term) and op.kind == "kick"):
                    raise _err(f"apply: {op.cls} is not supported (GridKick is)", ERR_INCOMPAT)
                self.items.append(_Kick(op, psi))
                continue
            if f.tag != "expeconly":
                raise _err(f"filter action <{f.tag}> is not supported", ERR_INCOMPAT)
            op = ld.op(f)
            fa = _attrs(f, srcdir)
            self.items.append(_Filter(ctx, op, psi, fa.get("header", "O" + op.cls)))
        if not self.items:
            raise _err("Empty filter list provided")
        shown = [x for x in self.items if isinstance(x, _Filter)]
        if shown:
            self.f = open(d + sa.get("fname", default_name), "w")
            self.f.write("time \t" + "".join(x.label + "\t" for x in shown) + "\n")

    def apply(self, psi, step, time):
        if not self.items:
            return None
        vals = []
        for x in self.items:
            if x is None:
                psi.normalize()
            elif isinstance(x, _Kick):
                x.apply(psi)
            else:
                v = x.value(psi, step)
                vals.append(v / psi.norm()[0] ** 2 if self.renorm else v)
        if self.f:
            self.f.write(_g10(time) + "\t" + "".join(_g10(v) + "\t" for v in vals) + "\n")
        return vals

    def close(self):
        if self.f:
            self.f.close()
            self.f = None


def hann(c):
    """FunctionGenerator::Hann (src/math/math_functions.h:108-113): the
    window over the whole record, w(0) = w(n-1) = 0."""
    c = np.asarray(c, dtype=complex)
    n = np.arange(c.size)
    return c * (0.5 * (1.0 - np.cos(2 * np.pi * n / (c.size - 1))))


def write_spectrum(corr, dt, path):
    """Reporter::WriteSpectrum (src/Reporter.cpp:160-190): angular frequency
    (i dw below n/2, -(n-i) dw above), |F|, Re F, Im F of the unnormalised
    forward transform divided by n/2 (integer), default stream format."""
    c = np.asarray(corr, dtype=complex)
    N = c.size
    F = np.fft.fft(c)
    dw = 2 * np.pi / (dt * N)
    h = N // 2
    with open(path, "w") as f:
        for i in range(N):
            w = i * dw if i < h else -(N - i) * dw
            f.write("%g\t%g\t%g\t%g\n" % (w, abs(F[i]) / h, F[i].real / h, F[i].imag / h))


# ---------------------------------------------------------------------------
# programs

class Result:
    """What a program produced: report rows (propa), energies (eigen), and
    the output folder."""

    def __init__(self, kind, directory):
        self.kind, self.dir = kind, directory
        self.report = []       # (step, time, norm2, pops, energy, autocorr)
        self.report_text = []
        self.expec_post = []   # (time, values)
        self.energies = None
        self.energies_raw = None
        self.steps = None


def _outdir(a, outdir):
    d = a.get("dir", "")
    if d and not d.endswith("/"):
        d += "/"
    if outdir:
        d = os.path.join(outdir, d)
    if d:
        os.makedirs(d, exist_ok=True)
    return d


def _report_line(step, time, n2, pops, energy, ac):
    s = "%d\t%.1f\t" % (step, time)
    s += "\t%.6f" % n2
    for p in pops:
        s += "\t%.6f" % p
    s += "\t%.8f" % energy
    s += "\t%.8f,%.8fi" % (ac.real, ac.imag)
    return s


def run_propa(ctx, el, srcdir=None, outdir=None, log=None):
    a = _attrs(el, srcdir)
    dt, steps = float(a.get("dt", 0)), int(a.get("steps", 0))
    if dt == 0:
        raise _err("Zero time step defined")
    if steps == 0:
        raise _err("Zero number of time step defined")
    wcycle = int(a.get("wcycle", 10))
    if wcycle < 1:
        raise _err("Write cycle less than one defined")
    # restart (ProgPropa.cpp:82-86, 292-314): the clock starts at step
    # `start`, the loop runs i = start .. steps (1 .. steps without it), the
    # initial state is not written and the file counter starts at `start`
    start = int(a.get("start", 0))
    if start < 0:
        raise _err("Only positive counter start counter numbers are allowed")
    d = _outdir(a, outdir)
    fname = a.get("fname", "WF")
    ld = Loader(srcdir, outdir)
    props = [c for c in el if c.tag == "propagator"]
    wfs = [c for c in el if c.tag == "wf"]
    if len(props) != 1 or len(wfs) != 1:
        raise _err("<propa> needs one <propagator> and one <wf>")
    spec = ld.propagator(props[0])
    psi = load_wave_function(ctx, wfs[0], srcdir, outdir)
    # initialisation filters, once, before psi0 is taken (ProgPropa.cpp:216-229)
    init = FilterList(ctx, ld, [c for c in el if c.tag == "filterinit"], psi, d, "expec_init.dat", srcdir,
                      renorm_default=False)
    init.apply(psi, 0, 0.0)
    init.close()
    U, H = build_propagator(ctx, spec, psi, dt)
    N, lo, hi = _grid_lists(psi)
    psi0 = WaveFunction(ctx, N, lo, hi, psi.S)
    psi0.copy_from(psi)
    pre = FilterList(ctx, ld, [c for c in el if c.tag == "filterpre"], psi, d, "expec_pre.dat", srcdir)
    post = FilterList(ctx, ld, [c for c in el if c.tag == "filterpost"], psi, d, "expec_post.dat", srcdir)
    res = Result("propa", d)
    out = open(d + a["nfile"], "w") if a.get("nfile") else None
    header = "Step\tTime[au]\tPopulation" + "".join("\t\tPop_%d" % s for s in range(psi.S) if psi.S > 1) + \
        "\t\tEnergy[au]\t<Psi0|PsiT>"
    emit = (lambda s: out.write(s + "\n")) if out else (log or (lambda s: None))
    emit(header)
    if start == 0:
        psi.save(d + fname, counter=0)
    nwritten = 1 if start == 0 else start
    spec = np.zeros(steps, dtype=complex) if "spectrum" in a else None
    c = start   # clock step
    try:
        for n_an, i in enumerate(range(max(start, 1), steps + 1)):
            pre.apply(psi, c, c * dt)
            # Reporter::Analyze on the state of clock step c
            n, per = psi.norm()
            pops = [float(x) for x in per] if psi.S > 1 else []   # squared norms per state
            e = H.expec(psi, c)
            ac = psi.overlap(psi0)
            row = (c, c * dt, n * n, pops, e, ac)
            if not math.isfinite(n):
                raise _err("Norm is not a number")
            res.report.append(row)
            if spec is not None:
                spec[n_an] = ac
            line = _report_line(*row)
            res.report_text.append(line)
            emit(line)
            U.apply(psi, c)
            c += 1
            vals = post.apply(psi, c, c * dt)
            if vals is not None:
                res.expec_post.append((c * dt, vals))
            if i % wcycle == 0:
                psi.save(d + fname, counter=nwritten)
                nwritten += 1
                write_propagation_meta(d, i, dt, fname, wcycle)
    finally:
        if out:
            out.close()
        pre.close()
        post.close()
    if spec is not None:   # Reporter::Finalize (src/Reporter.cpp:199-215)
        write_spectrum(hann(spec) if _bool(a.get("win"), True) else spec, dt, d + a["spectrum"])
    res.psi = psi
    return res


def run_eigen(ctx, el, srcdir=None, outdir=None, log=None):
    a = _attrs(el, srcdir)
    method = a.get("method", "imag")
    if method not in ("imag", "ac"):
        raise _err(f"eigen method '{method}' is not supported", ERR_INCOMPAT)
    if "read" in a:
        raise _err("reading a stored propagation (read=) is not supported", ERR_INCOMPAT)
    if "dt" not in a:
        raise _err("No time step defined")
    dt = float(a["dt"])
    nef = int(a.get("Nef", 20))
    # ProgEigen.cpp:187, 206: diag defaults on for imag, off for ac
    diag = _bool(a.get("diag"), method == "imag")
    conv = float(a["conv"]) if "conv" in a else None
    max_steps = int(a.get("steps", 512))
    d = _outdir(a, outdir)
    ld = Loader(srcdir, outdir)
    props = [c for c in el if c.tag == "propagator"]
    wfs = [c for c in el if c.tag == "wf"]
    if len(props) != 1 or len(wfs) != 1:
        raise _err("<eigen> needs one <propagator> and one <wf>")
    spec = ld.propagator(props[0])
    psi = load_wave_function(ctx, wfs[0], srcdir, outdir)
    st = ac = None
    if method == "imag":
        U, H = build_propagator(ctx, spec, psi, dt, imag=True)
        efs, e, er, st = eigen_imag(U, psi, nef, max_steps, conv, diag, H)
    else:   # ProgEigen::AutoCorrEF: real time, one function per spectral peak
        U, H = build_propagator(ctx, spec, psi, dt)
        win = _bool(a.get("win"), True)
        efs, e, er, ac = eigen_ac(U, psi, max_steps, win, float(a.get("tol", 1.0)), int(a.get("width", -1)),
                                  diag, H)
        nef = len(efs)
        write_spectrum(hann(ac) if win else ac, dt, d + a.get("spectrum", "spectrum.dat"))
    fname = a.get("fname", "ef")
    for k, w in enumerate(efs):
        w.save(d + fname, counter=k)
    en = e if diag else er
    with open(d + a.get("ename", "energy.dat"), "w") as f:
        for i in range(nef):
            f.write("%d\t%.8f" % (i, en[i]))
            f.write("\t%.2f" % (en[i] * AU2WAVENUMBERS))
            f.write("\t%.2f" % ((en[i] - en[0]) * AU2WAVENUMBERS))
            if i > 0:
                f.write("\t%.2f" % ((en[i] - en[i - 1]) * AU2WAVENUMBERS))
            f.write("\n")
    res = Result("eigen", d)
    res.energies, res.energies_raw, res.steps, res.efs, res.autocorr = e, er, st, efs, ac
    if log:
        for i in range(nef):
            log("%d\t%.8f" % (i, en[i]))
    return res


def _projection_target(ctx, op, psi, srcdir, base):
    """An OCT operator target as a device projection: a Projection
    (single state), or a Multistate of Projections on diagonal elements
    with unity off (each function placed in its state, zero elsewhere)."""
    N, lo, hi = _grid_lists(psi)

    def efs(t):
        p = t.params
        return [WaveFunction.load(ctx, _resolve(p["files"], base), counter=p["start"] + k * p["step"])
                for k in range(p["num"])]
    if isinstance(op, Term) and op.kind == "proj":
        if psi.S != 1:
            raise _err("a Projection target of a multistate function: use a Multistate target", ERR_INCOMPAT)
        return Projection(efs(op))
    if isinstance(op, Multi):
        if op.unity:
            raise _err("a Multistate target needs unity=\"false\"", ERR_INCOMPAT)
        wfs = []
        for (i, j), e in sorted(op.elems.items()):
            if i != j or not (isinstance(e, Term) and e.kind == "proj"):
                raise _err("Multistate targets hold Projections on the diagonal only", ERR_INCOMPAT)
            for f in efs(e):
                w = WaveFunction(ctx, N, lo, hi, psi.S)
                w.upload(np.zeros(psi.S * w.size, dtype=np.complex128))
                w.upload(f.download(), i)
                wfs.append(w)
        return Projection(wfs)
    raise _err("OCT operator targets: Projection or a Multistate of Projections", ERR_INCOMPAT)


def _oct_table(rows, nt, overlap, phase):
    """ProgOCT::Report's iteration table (ProgOCT.cpp:363-425): fixed
    notation with 8 decimals."""
    head = "iteration\t"
    for t in range(nt):
        head += "Norm_%d\t\t" % t
        if overlap:
            head += "Norm_t%d\t\tOverlapp_%d\t" % (t, t)
        else:
            head += "Operator_%d\t" % t
        if phase and overlap:
            head += "Phase_%d\t\t" % t
    head += "Laser Objective\t\tObjective\t\tPulse Energy"
    lines = [head]
    for r in rows:
        it = int(r[0])
        s = "%d\t\t" % it
        for t in range(nt):
            b = 5 + 5 * t
            s += "%.8f\t" % r[b]
            if overlap:
                s += "%.8f\t%.8f\t" % (r[b + 1], r[b + 2])
                if phase:
                    s += "%.8f\t" % r[b + 3]
            else:
                s += "%.8f\t" % r[b + 4]
        s += "-\t\t\t-\t\t\t" if it == 0 else "%.8f\t\t%.8f\t\t" % (r[1], r[2])
        s += "%.8f" % r[3]
        lines.append(s)
    return lines


def run_oct(ctx, el, srcdir=None, outdir=None, log=None):
    """ProgOCT (src/ProgOCT.cpp:59-300, 868-1092) on the device."""
    a = _attrs(el, srcdir)
    dt, steps = float(a.get("dt", 0)), int(a.get("steps", 0))
    if dt == 0:
        raise _err("Zero time step defined")
    if steps == 0:
        raise _err("Zero number of time step defined")
    method = a.get("method", "krotov")
    if method == "freq":
        raise _err("OCT method 'freq' (frequency shaping) is not supported", ERR_INCOMPAT)
    if method not in ("krotov", "rabitz", "rabitzfb"):
        raise _err(f"Unknown OCT method {method}")
    if a.get("coup", "dipole") != "dipole":
        raise _err("Unsupported coupling specified")
    ttype = a.get("ttype", "overlap")
    if ttype not in ("overlap", "operator"):
        raise _err(f"Unknown OCT target type {ttype}")
    if _bool(a.get("mv")):
        raise _err("moving targets (mv) are not supported", ERR_INCOMPAT)
    if _bool(a.get("writel")):
        raise _err("writing the laser of every iteration (writel) is not supported", ERR_INCOMPAT)
    if "shape" not in a:
        raise _err("No shape file given")
    iterations = int(a.get("iterations", 50))
    conv = float(a.get("conv", 1e-5))
    alpha = float(a.get("alpha", 1.0))
    nt = int(a.get("ntargets", 1))
    phase = _bool(a.get("phase"))
    d = _outdir(a, outdir)
    ld = Loader(srcdir, outdir)
    props = [c for c in el if c.tag == "propagator"]
    if len(props) != 1:
        raise _err("No propagator found")
    spec = ld.propagator(props[0])
    sect = {c.tag: c for c in el}
    psi_i = []
    for t in range(nt):
        if "wfi%d" % t not in sect:
            raise _err("Missing initial wavefunction %d" % t)
        psi_i.append(load_wave_function(ctx, sect["wfi%d" % t], srcdir, outdir))
    U, _ = build_propagator(ctx, spec, psi_i[0], dt)
    kw = {}
    if ttype == "overlap":
        psi_t = []
        for t in range(nt):
            if "wft%d" % t not in sect:
                raise _err("Missing target wavefunction %d" % t)
            psi_t.append(load_wave_function(ctx, sect["wft%d" % t], srcdir, outdir))
        kw["psi_t"] = psi_t
    else:
        tg = []
        for t in range(nt):
            if "target%d" % t not in sect:
                raise _err("Missing target operator %d" % t)
            tg.append(_projection_target(ctx, ld.op(sect["target%d" % t]), psi_i[0], srcdir, outdir))
        kw["targets"] = tg
    G = None
    if "gobbler" in sect:
        G = build_hamiltonian(ctx, ld.op(sect["gobbler"]), psi_i[0])
    _, shape = load_laser(_resolve(a["shape"], outdir))
    # the laser carried by the propagator's dipole terms and its file's dt
    H = U.H if hasattr(U, "H") else next(o for o in U.ops if _has_dipole(o))
    laser_dt = _dipole_laser_dt(spec)
    rows, converged = oct_run(U, shape, psi_i, steps, method=method, phase=phase, iterations=iterations,
                              conv=conv, alpha=alpha, gobbler=G, **kw)
    save_laser(d + a.get("fname", "laser"), laser_dt, H.laser())
    lines = _oct_table(rows, nt, ttype == "overlap", phase)
    if _bool(a.get("writeyield")) or "yfile" in a:
        with open(d + (a.get("yfile") or "Yield"), "w") as f:
            f.write("\n".join(lines[1:]) + "\n")
    if log:
        for ln in lines:
            log(ln)
        if converged:
            log("Convergence reached")
    res = Result("oct", d)
    res.rows, res.converged, res.table = rows, converged, lines
    return res


def _has_dipole(H):
    try:
        H.laser()
        return True
    except QdxError:
        return False


def _dipole_laser_dt(spec):
    """dt of the first GridDipole's laser file in the propagator's operators."""
    def walk(o):
        if isinstance(o, Term):
            if o.kind == "dip":
                yield o
        elif isinstance(o, Sum):
            for c in o.items:
                yield from walk(c)
        elif isinstance(o, Multi):
            for c in o.elems.values():
                yield from walk(c)
    for op in spec[2]:
        for t in walk(op):
            return load_laser(t.params["laser"])[0]
    raise _err("Given type of coupling operator not found in hamiltonian")


PROGRAMS = {"propa": run_propa, "eigen": run_eigen, "oct": run_oct}


def programs(path):
    """The program elements of a .qdi file, in order."""
    root = ET.parse(path).getroot()
    if root.tag != "qdng":
        raise _err(f"{path}: the root element is <{root.tag}>, not <qdng>")
    return list(root)


def run(path, ctx, srcdir=None, outdir=None, which=None, log=None):
    """Run the programs of a QDng input file (all, or the indices in `which`)
    on the device.  srcdir replaces @srcdir@ in attribute values; outdir
    prefixes the programs' `dir`.  The root element's attributes are QDng's
    global parameters (CmdHandler.cpp:97); of them wf_format, compress and
    comptol (FileWF.cpp:17-56) set the output file format for the run.
    Returns one Result per program run."""
    root = ET.parse(path).getroot()
    if root.tag != "qdng":
        raise _err(f"{path}: the root element is <{root.tag}>, not <qdng>")
    old = get_file_format()
    g = root.attrib
    fmt = g.get("wf_format", old[0])
    if fmt not in ("binary", "stream"):
        raise _err(f"Unrecognized file wave function file format: {fmt}")
    out = []
    try:
        set_file_format(fmt, _bool(g.get("compress"), old[1]), float(g.get("comptol", old[2])))
        for k, el in enumerate(list(root)):
            if which is not None and k not in which:
                continue
            if el.tag not in PROGRAMS:
                raise _err(f"program <{el.tag}> is not supported", ERR_INCOMPAT)
            out.append(PROGRAMS[el.tag](ctx, el, srcdir, outdir, log))
    finally:
        set_file_format(*old)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run a QDng program input file on the GPU")
    ap.add_argument("input")
    ap.add_argument("-d", "--dir", default=None, help="output folder prefix")
    ap.add_argument("--srcdir", default=None, help="value of @srcdir@")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    from . import Context
    ctx = Context(args.device)
    try:
        run(args.input, ctx, args.srcdir, args.dir, log=lambda s: print(s, flush=True))
    finally:
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
