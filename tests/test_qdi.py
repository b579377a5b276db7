"""QDng program inputs (XML .qdi) through qdng_amd.qdi.

The inputs are the programs of QDng's grid regression cases, generated from
their parameters by tests/golden/qdi_programs.py; the data files they read
(potentials, dipoles, lasers, wave functions) are in tests/golden/qdng_files.  CPU
tests check the operator graphs the loader builds (key/ref sharing, matrix
elements, @srcdir@) and its error paths; the `gpu` tests run the programs end
to end and compare the last report line and the last expec_post.dat line with
the values tests/checktests.sh greps for, and the eigen programs' energy.dat
with CheckTGE-im-Energy (checktests.sh:4-29)."""
import os
import sys
import tempfile

import numpy as np
import pytest

import cases
from oracle import oracle as O
from test_oracle import _digits_ok
import qdng_amd as Q
from qdng_amd import qdi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = os.path.join(GOLDEN, "qdng_files")
sys.path.insert(0, GOLDEN)
import qdi_programs  # noqa: E402

_PROGRAMS = qdi_programs.write_all(tempfile.mkdtemp(prefix="qdi_programs_"))
PROPA = ["TGP-Expec", "TGPMS-Cheby", "TGPMS-GSPO", "TGPL-IR-Pulse-Cheby", "TGPL-IR-Pulse-GSPO",
         "TGPLMS-Cheby", "TGPLMS-SIL", "TGP-Kick", "TGP-NIP-Cheby"]
# checktests.sh: CheckT TGP-Kick and CheckT-Expec TGP-Kick (grep prefixes)
KICK = {"report": {"step": 999, "time": 19980.0, "population": 1.0, "pops": [], "energy": 0.00399372,
                   "autocorr": [0.29628786, -0.88631099]},
        "expec_post": {"time": 20000.0, "prefix": [True] * 5,
                       "values": ["-0.81968278", "3.23866603", "0.99999811", "0.00195512887", "0.00203859515"]}}
EIGEN = ["TGE-Cheby", "TGE-GSPO", "TGE-SIL", "TGE-Arnoldi"]
# checktests.sh:31-45, CheckTGE-ac-Energy
TGE_AC_ENERGIES = ["0.003714", "0.011074", "0.018152", "0.024892", "0.031319", "0.037427"]
# checktests.sh: CheckT-Expec TGPLMS-SIL (the report line equals TGPLMS-Cheby's)
SIL_POST = ["-0.3059539159", "3.364408943", "1.661276151", "1.739725823", "0.001196532131", "0.007276466639",
            "0.001079240961", "0.07679400487"]
TGE_ENERGIES = ["0.0037100", "0.0110668", "0.0181411", "0.0249017", "0.0313281", "0.0374320",
                "0.0431786", "0.0485", "0.0535", "0.058"]


def qdi_path(name):
    return _PROGRAMS[name]


def prog(name, k=0):
    return qdi.programs(qdi_path(name))[k]


def test_every_input_parses():
    for name in PROPA + EIGEN:
        for el in qdi.programs(qdi_path(name)):
            assert el.tag in ("propa", "eigen")
            ld = qdi.Loader(FIX)
            props = [c for c in el if c.tag == "propagator"]
            assert len(props) == 1
            pname, attrs, ops = ld.propagator(props[0])
            assert pname in ("Cheby", "GSPO", "SIL", "Arnoldi")
            for sec in (c for c in el if c.tag == "filterpost"):
                for f in sec:
                    ld.op(f)


def test_gobbler_terms():
    ld = qdi.Loader(FIX)
    el = prog("TGP-NIP-Cheby")
    name, a, (H,) = ld.propagator([c for c in el if c.tag == "propagator"][0])
    assert name == "Cheby" and "prec" not in a
    kin, nip, wall = qdi._terms(H)
    assert kin.kind == "kin" and nip.kind == "gob" and wall.kind == "gob"
    assert nip.params == dict(lp={}, rp={0: 5.5}, order=5, gain=0.02, nip=True, dims=1)
    assert wall.params == dict(lp={0: 2.0}, rp={}, order=20, gain=0.3, nip=False, dims=1)
    assert ld.keys["V"] is wall and ld.keys["T"] is kin
    # defaults (OGobbler::Init): order 20, gain 1, nip off
    g = ld.op(ET_fromstring('<G name="Gobbler" lp1="0.5"/>'))
    assert g.params == dict(lp={1: 0.5}, rp={}, order=20, gain=1.0, nip=False, dims=None)


def ET_fromstring(x):
    import xml.etree.ElementTree as ET
    return ET.fromstring(x)


def test_multistate_graph_with_refs():
    ld = qdi.Loader(FIX)
    el = prog("TGPLMS-Cheby")
    name, a, (H,) = ld.propagator([c for c in el if c.tag == "propagator"][0])
    assert name == "Cheby" and float(a["prec"]) == 1e-12
    assert isinstance(H, qdi.Multi) and not H.dmulti and H.unity and H.S == 2
    assert sorted(H.elems) == [(0, 0), (1, 0), (1, 1)]
    t0, v0 = H.elems[(0, 0)].items
    t1, v1 = H.elems[(1, 1)].items
    assert t0 is t1 and t0.kind == "kin" and t0.params["mass"] == [1763.2]   # <Tkin ref="T"/>
    assert v0.params["file"] == os.path.join(FIX, "GridPot0") and v1.params["file"].endswith("GridPot1")
    dip = H.elems[(1, 0)]
    assert dip.kind == "dip" and dip.params["laser"] == os.path.join(FIX, "GridLaser_w0.176")
    # the filters refer to the Hamiltonian's terms by key
    filt = [ld.op(f) for f in [c for c in el if c.tag == "filterpost"][0]]
    assert [type(f) for f in filt] == [qdi.Multi] * 8
    assert filt[1].elems[(1, 1)] is filt[0].elems[(0, 0)]               # <m1.1 ref="p"/>
    assert filt[6].elems[(0, 0)] is v0 and filt[7].elems[(1, 1)] is v1
    assert not filt[0].unity and filt[0].S == 2


def test_gspo_split_operators_and_their_sum():
    ld = qdi.Loader(FIX)
    el = prog("TGPMS-GSPO")
    name, a, ops = ld.propagator([c for c in el if c.tag == "propagator"][0])
    assert name == "GSPO" and len(ops) == 2 and all(o.dmulti for o in ops)
    m = qdi._merged(ops)
    assert sorted(m.elems) == [(0, 0), (1, 1)]
    assert [t.kind for t in qdi._terms(m.elems[(1, 1)])] == ["kin", "pot"]
    ld = qdi.Loader(FIX)
    el = prog("TGPL-IR-Pulse-GSPO")
    name, a, ops = ld.propagator([c for c in el if c.tag == "propagator"][0])
    assert [o.kind for o in ops] == ["kin", "pot", "dip"]
    assert [t.kind for t in qdi._terms(qdi._merged(ops))] == ["kin", "pot", "dip"]


def test_loader_errors(tmp_path):
    import xml.etree.ElementTree as ET
    ld = qdi.Loader()
    for xml, code in [('<A ref="nope"/>', -1), ('<A name="Flux"/>', -2), ('<A name="GridKick"/>', -1), ('<A/>', -1),
                      ('<A name="Gobbler" order="0"/>', -1), ('<A name="Gobbler" residue="true"/>', -2),
                      ('<A name="GridPotential"/>', -1), ('<A name="GridDipole" file="x"/>', -1),
                      ('<A name="Multistate"><x0 name="GridNablaSq" mass="1"/></A>', -1),
                      ('<A name="Sum"/>', -1), ('<A name="GridPotential" file="f" setzero="maybe"/>', -1)]:
        with pytest.raises(Q.QdxError) as e:
            ld.op(ET.fromstring(xml))
        assert e.value.code == code, xml
    with pytest.raises(Q.QdxError) as e:
        ld.propagator(ET.fromstring('<propagator name="SPO"/>'))
    assert e.value.code == -2
    with pytest.raises(Q.QdxError):
        ld.propagator(ET.fromstring('<propagator name="GSPO"><A name="GridNablaSq" mass="1"/></propagator>'))
    # a kinetic operator off the diagonal, a filter inside a Hamiltonian
    m = ld.op(ET.fromstring('<H name="Multistate"><m1.0 name="GridNablaSq" mass="1"/></H>'))
    with pytest.raises(Q.QdxError):
        qdi._add_term(None, 1, 0, m.elems[(1, 0)])
    with pytest.raises(Q.QdxError) as e:
        qdi._terms(ld.op(ET.fromstring('<F name="GridMomentum"/>')))
    assert e.value.code == -2
    # unknown eigen methods and stored propagations are rejected before
    # anything touches a device
    for attrs in ('method="lanczos"', 'method="ac" read="TGE-run"'):
        f = tmp_path / "e.qdi"
        f.write_text('<qdng><eigen dt="5" %s/></qdng>' % attrs)
        with pytest.raises(Q.QdxError) as e:
            qdi.run(str(f), None, FIX, str(tmp_path))
        assert e.value.code == -2
    bad = tmp_path / "bad.qdi"
    bad.write_text("<root><propa/></root>")
    with pytest.raises(Q.QdxError):
        qdi.programs(str(bad))


# ---------------------------------------------------------------------------
# end to end on the device

@pytest.mark.gpu
@pytest.mark.parametrize("name", PROPA)
def test_propagation_inputs(ctx, name, tmp_path):
    print("[gpu case] qdi", name, flush=True)
    (res,) = qdi.run(qdi_path(name), ctx, FIX, str(tmp_path))
    gold = KICK if name == "TGP-Kick" else \
        cases.kat_data()[1]["golden"]["TGPLMS-Cheby" if name == "TGPLMS-SIL" else name]
    rep = gold["report"]
    step, time, n2, pops, energy, ac = res.report[-1]
    assert (step, time) == (rep["step"], rep["time"])
    assert abs(n2 - rep["population"]) < 5e-7
    for got, want in zip(pops, rep["pops"]):
        assert abs(got - want) < 5e-7
    assert abs(energy - rep["energy"]) < 6e-9, (energy, rep["energy"])
    assert abs(ac.real - rep["autocorr"][0]) < 6e-9 and abs(ac.imag - rep["autocorr"][1]) < 6e-9
    # expec_post.dat: header plus one line per step, the last at steps * dt
    lines = open(os.path.join(res.dir, "expec_post.dat")).read().splitlines()
    assert lines[0].startswith("time \t") and len(lines) == 1 + len(res.report)
    last = lines[-1].split("\t")
    post = gold["expec_post"]
    assert float(last[0]) == pytest.approx(post["time"])
    want = SIL_POST if name == "TGPLMS-SIL" else post["values"]
    prefix = [False] * len(SIL_POST) if name == "TGPLMS-SIL" else post["prefix"]
    vals = [float(v) for v in last[1:] if v]
    assert len(vals) == len(want)
    for v, w, pf in zip(vals, want, prefix):
        assert _digits_ok(v, w, pf), (name, v, w)
    # spectrum.dat: one line per report line; the ground-state inputs put the
    # largest peak on the energy (<psi(t)|psi0> ~ e^{+iEt})
    spec = [[float(v) for v in l.split("\t")] for l in open(os.path.join(res.dir, "spectrum.dat"))]
    assert len(spec) == len(res.report)
    if name == "TGP-Expec":
        top = max(spec, key=lambda r: r[1])
        assert abs(top[0] - energy) < 2 * spec[1][0]   # spec[1][0] = dw
    # wave functions: the initial one and one per write cycle, Propagation.meta
    a = open(qdi_path(name)).read()
    steps = len(res.report)
    wcycle = int(a.split('wcycle="')[1].split('"')[0])
    nfiles = 1 + steps // wcycle
    base = os.path.join(res.dir, "WF")
    S = res.psi.S
    for k in range(nfiles):
        f = base + ("-0" if S > 1 else "") + "_%d.wf" % k
        assert os.path.exists(f), f
    if steps >= wcycle:
        assert Q.read_meta(os.path.join(res.dir, "Propagation.meta"))["Nt"] == str(wcycle * (steps // wcycle))
        last_wf = Q.WaveFunction.load(ctx, base, counter=nfiles - 1)
        assert np.array_equal(last_wf.download(), res.psi.download())


@pytest.mark.gpu
@pytest.mark.parametrize("name", EIGEN)
def test_eigen_inputs(ctx, name, tmp_path):
    print("[gpu case] qdi", name, flush=True)
    res, res_ac = qdi.run(qdi_path(name), ctx, FIX, str(tmp_path))
    # method "ac" on the LC of the first six functions the imag program wrote
    rows = [l.split("\t") for l in open(os.path.join(res_ac.dir, "energy.dat")).read().splitlines()]
    assert [r[0] for r in rows] == [str(i) for i in range(len(rows))] and len(rows) >= 6
    for r, want in zip(rows, TGE_AC_ENERGIES):
        assert r[1].startswith(want), (name, "ac", r[1], want)
    assert os.path.exists(os.path.join(res_ac.dir, "ef_5.wf"))
    assert len(open(os.path.join(res_ac.dir, "spectrum.dat")).read().splitlines()) == 2048
    rows = [l.split("\t") for l in open(os.path.join(res.dir, "energy.dat")).read().splitlines()]
    assert [r[0] for r in rows] == [str(i) for i in range(10)]
    for r, want in zip(rows, TGE_ENERGIES):
        assert r[1].startswith(want), (name, r[1], want)
    assert float(rows[0][2]) == pytest.approx(float(rows[0][1]) * 219475, abs=0.01)
    for k in range(10):
        assert os.path.exists(os.path.join(res.dir, "ef_%d.wf" % k))
    ef = Q.WaveFunction.load(ctx, os.path.join(res.dir, "ef"), counter=3)
    assert ef.norm()[0] == pytest.approx(1.0, abs=1e-9)


def test_spectrum_file(tmp_path):
    # a pure tone on bin 3 of 16: amplitude N/(N/2) = 2 at w = 3 dw, zero elsewhere
    N, dt = 16, 0.5
    dw = 2 * np.pi / (dt * N)
    c = np.exp(1j * 3 * dw * dt * np.arange(N))
    qdi.write_spectrum(c, dt, str(tmp_path / "s.dat"))
    rows = [[float(v) for v in l.split("\t")] for l in open(tmp_path / "s.dat").read().splitlines()]
    assert len(rows) == N
    assert [r[0] for r in rows] == pytest.approx([i * dw if i < N // 2 else -(N - i) * dw for i in range(N)],
                                                  rel=1e-5)
    assert rows[3][1] == pytest.approx(2.0, rel=1e-6) and rows[3][2] == pytest.approx(2.0, rel=1e-6)
    assert max(abs(r[1]) for i, r in enumerate(rows) if i != 3) < 1e-9
    # the Hann window zeroes both ends and is symmetric
    w = qdi.hann(np.ones(9))
    assert w[0] == 0 and abs(w[-1]) < 1e-15 and w[4] == pytest.approx(1.0)
    assert np.allclose(w, w[::-1])


def test_report_line_format():
    # Reporter::Analyze's line (src/Reporter.cpp:92-150) for the values of
    # checktests.sh's TGP-Expec and TGPMS-Cheby lines, compared token by token
    got = qdi._report_line(999, 19980.0, 1.0, [], 0.00371015, complex(0.29455816, -0.95562855))
    assert got.split() == "999    19980.0         1.000000        0.00371015      0.29455816,-0.95562855i".split()
    got = qdi._report_line(999, 9990.0, 1.0, [0.5, 0.5], 0.08955451, complex(0.43841526, -0.36254206))
    want = "999    9990.0          1.000000        0.500000        0.500000        0.08955451      0.43841526,-0.36254206i"
    assert got.split() == want.split()
    assert got.startswith("999\t9990.0\t\t1.000000\t0.500000")


@pytest.mark.gpu
def test_pre_filters_normalize_and_renorm(ctx, tmp_path):
    # TGP-Expec shortened to 20 steps, with a pre-step section (position,
    # then normalize) and a post section without renormalisation
    from oracle import oracle as O
    print("[gpu case] qdi filters", flush=True)
    src = open(qdi_path("TGP-Expec")).read()
    src = src.replace('steps="1000" wcycle="1000"', 'steps="20" wcycle="10"')
    src = src.replace("<filterpost>", '<filterpre fname="pre.dat"><expeconly name="GridPosition"/><normalize/>'
                                      '</filterpre><filterpost renorm="false">')
    f = tmp_path / "in.qdi"
    f.write_text(src)
    (res,) = qdi.run(str(f), ctx, FIX, str(tmp_path))
    pre = [l.split("\t") for l in open(os.path.join(res.dir, "pre.dat")).read().splitlines()]
    assert pre[0] == ["time ", "OGridPosition", ""] and len(pre) == 21
    assert [float(r[0]) for r in pre[1:3]] == [0.0, 20.0]
    # the first pre-step value is <x> of the normalised initial state
    g = O.Grid([128], [1.8897], [6.1104])
    psi0 = np.fromfile(os.path.join(FIX, "GridPot0_ef_0.wf"), dtype="<c16")
    psi0 = psi0 / np.sqrt(O.norm2(g, psi0))
    assert float(pre[1][1]) == pytest.approx(O.expec_position(g, psi0), rel=1e-9)
    post = open(os.path.join(res.dir, "expec_post.dat")).read().splitlines()
    assert len(post) == 21 and len(res.expec_post) == 20
    # files at steps 10 and 20 besides the initial one
    assert os.path.exists(os.path.join(res.dir, "WF_2.wf"))
    assert Q.read_meta(os.path.join(res.dir, "Propagation.meta"))["Nt"] == "20"


def test_filter_renorm_defaults(tmp_path, monkeypatch):
    # FiltersList starts with renorm off (src/FiltersList.cpp:11); ProgPropa
    # switches it on for the pre and post sections only (ProgPropa.cpp:260,
    # 276), so an expeconly in <filterinit> reports <psi|O|psi> undivided
    # for an unnormalised psi; the section's renorm attribute overrides both
    import xml.etree.ElementTree as ET

    class FakeFilter:
        def __init__(self, ctx, op, psi, label):
            self.label = label

        def value(self, psi, step):
            return 3.0

    class FakePsi:
        def norm(self):
            return (2.0, [2.0])          # <psi|psi> = 4

    monkeypatch.setattr(qdi, "_Filter", FakeFilter)
    ld = qdi.Loader(FIX)
    d = str(tmp_path) + "/"

    def values(xml, **kw):
        fl = qdi.FilterList(None, ld, [ET.fromstring(xml)], FakePsi(), d, "expec.dat", FIX, **kw)
        v = fl.apply(FakePsi(), 0, 0.0)
        fl.close()
        return v

    sec = '<filterinit%s><expeconly name="GridPosition"/></filterinit>'
    assert values(sec % "", renorm_default=False) == [3.0]
    assert values(sec % ' renorm="true"', renorm_default=False) == [0.75]
    assert values(sec.replace("init", "post") % "") == [0.75]
    assert values(sec.replace("init", "post") % ' renorm="false"') == [3.0]


# ---------------------------------------------------------------------------
# optimal control programs (ProgOCT)

OCT = ["TGO-IR-Krotov", "TGO-IR-Rabitz", "TGO-IR-Krotov-Projection", "TGO-UV-Krotov"]


def oct_program(name, tmp_path, iterations):
    """The generated program with fewer iterations."""
    import xml.etree.ElementTree as ET
    tree = ET.parse(qdi_path(name))
    tree.getroot().find("oct").set("iterations", str(iterations))
    path = str(tmp_path / (name + ".qdi"))
    tree.write(path)
    return path


def test_oct_inputs_parse_and_reject(tmp_path):
    for name in OCT:
        els = qdi.programs(qdi_path(name))
        assert els[0].tag == "oct"
        ld = qdi.Loader(FIX)
        pname, _, _ = ld.propagator(els[0].find("propagator"))
        assert pname == "Cheby"
    # unsupported options fail before anything touches a device
    for attrs, code in (('method="freq"', -2), ('mv="true"', -2), ('writel="true"', -2), ('method="x"', -1),
                        ('ttype="phase"', -1), ('coup="raman"', -1), ('dt="0"', -1)):
        f = tmp_path / "o.qdi"
        base = {"dt": "1", "steps": "10", "shape": "s"}
        xml = "<qdng><oct " + " ".join('%s="%s"' % kv for kv in base.items() if kv[0] + "=" not in attrs)
        f.write_text(xml + " " + attrs + "/></qdng>")
        with pytest.raises(Q.QdxError) as e:
            qdi.run(str(f), None, FIX, str(tmp_path))
        assert e.value.code == code, attrs


def _oracle_uv_krotov(iterations):
    """TGO-UV-Krotov restated on the oracle: the TGPLMS Hamiltonian (two
    states, GridDip01 on m1.0 with GridLaser_w0.176), psi_i the ground state
    on state 0, the projection onto GridPot0_ef_0..3 in state 1."""
    a, _ = cases.kat_data()
    p, _, dt, steps, _ = cases.kat_problem("TGPLMS-Cheby")
    O.reset()
    H = p.oracle_ham()
    U = O.Cheby(H, dt, prec=1e-12)
    g = p.oracle_grid()
    coup = O.multistate(2, {(1, 0): O.potential(g, a["GridDip01"])}, unity=False)
    n = p.size
    vecs = []
    for k in range(4):
        v = np.zeros(2 * n, np.complex128)
        v[n:] = a["GridPot0_ef_%d" % k]
        vecs.append(v)
    _, shape = Q.load_laser(os.path.join(FIX, "GridShapeUV"))
    return O.oct_run(U, p.oracle_dipoles, a["GridLaser_w0.176"], coup, g, 2, shape, [p.psi0.copy()], steps, dt,
                     targets=[(vecs, True, False)], method="krotov", iterations=iterations, alpha=10.0,
                     conv=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("name", OCT)
def test_oct_inputs(ctx, name, tmp_path):
    print("[gpu case] qdi oct", name, flush=True)
    its = 3 if name == "TGO-UV-Krotov" else 2
    res = qdi.run(oct_program(name, tmp_path, its), ctx, FIX, str(tmp_path))
    o = res[0]
    assert o.kind == "oct" and len(o.rows) == its + 1
    y = o.rows[:, 4]
    if "Krotov" in name:
        assert np.all(np.diff(y) > 0), y      # Krotov increases the yield
    assert len(o.table) == its + 2 and o.table[0].startswith("iteration\t")
    dt_l, field = Q.load_laser(os.path.join(o.dir, "laser"))
    assert dt_l == (1.0 if name == "TGO-UV-Krotov" else 10.0)
    assert np.any(field != Q.load_laser(os.path.join(FIX, "GridLaser_w0.176" if name == "TGO-UV-Krotov"
                                                     else "GridLaser_w0.00713"))[1])
    if name == "TGO-UV-Krotov":
        rows_o, field_o, _ = _oracle_uv_krotov(its)
        m = ~np.isnan(rows_o)
        assert np.array_equal(np.isnan(o.rows), np.isnan(rows_o))
        assert np.max(np.abs(o.rows[m] - rows_o[m]) / np.maximum(np.abs(rows_o[m]), 1e-12)) < 1e-8
        assert np.max(np.abs(field - field_o)) < 1e-9
    else:
        # the propagation that follows reads the optimised laser (GSPO): its
        # final |<ef_1|psi>|^2 is the last OCT yield (Cheby), to the
        # difference of the two propagators
        propa = res[1]
        assert propa.kind == "propa"
        last = open(os.path.join(propa.dir, "expec_post.dat")).read().splitlines()[-1].split("\t")
        vals = [float(v) for v in last[1:] if v]
        pop1 = vals[3]                        # momentum, position, projections 0, 1, 2, T, V
        want = o.rows[-1, 5 + 2] if "Projection" not in name else o.rows[-1, 5 + 4]
        assert abs(pop1 - want) < 0.02 * want + 1e-4, (pop1, want)


@pytest.mark.gpu
def test_global_file_format_parameters(ctx, tmp_path):
    # <qdng wf_format="stream" compress="true">: the propagation's wave
    # functions are written as compressed streams, and the format is reset
    # after the run
    import xml.etree.ElementTree as ET
    tree = ET.parse(qdi_path("TGP-Expec"))
    root = tree.getroot()
    root.set("wf_format", "stream")
    root.set("compress", "true")
    pr = root.find("propa")
    pr.set("steps", "20")
    pr.set("wcycle", "10")
    path = str(tmp_path / "in.qdi")
    tree.write(path)
    (res,) = qdi.run(path, ctx, FIX, str(tmp_path))
    f = os.path.join(res.dir, "WF_2.wf")
    assert Q.file_format(f) == "stream"
    assert Q.get_file_format() == ("binary", False, 1e-6)
    w = Q.WaveFunction.load(ctx, os.path.join(res.dir, "WF"), counter=2)
    ref = res.psi.download()
    assert np.max(np.abs(w.download() - ref)) < 1e-4 * np.max(np.abs(ref))


@pytest.mark.gpu
def test_propagation_restart(ctx, tmp_path):
    # ProgPropa restart (src/ProgPropa.cpp:82-86, 292-326): a run from the
    # state written after 20 steps with start="20" and steps="30" takes steps
    # 20 .. 30 of the clock (11 steps, laser samples 20 .. 30) and ends where
    # an uninterrupted 31-step run ends; its report rows are that run's rows
    # 20 .. 30 and its files are numbered from 20
    import xml.etree.ElementTree as ET

    def run(tag, steps, start=None, wf=None):
        print("[gpu case] qdi restart", tag, flush=True)
        tree = ET.parse(qdi_path("TGPL-IR-Pulse-Cheby"))
        pr = tree.getroot().find("propa")
        pr.set("steps", str(steps))
        pr.set("wcycle", "10")
        pr.set("dir", tag)
        if start is not None:
            pr.set("start", str(start))
            w = pr.find("wf")
            w.set("file", wf)
            del w.attrib["normalize"]
        path = str(tmp_path / (tag + ".qdi"))
        tree.write(path)
        (res,) = qdi.run(path, ctx, FIX, str(tmp_path))
        return res

    first = run("first", 20)
    full = run("full", 31)
    rs = run("restart", 30, 20, os.path.join(first.dir, "WF_2"))
    assert [r[:2] for r in rs.report] == [r[:2] for r in full.report[20:31]]
    for a, b in zip(rs.report, full.report[20:31]):
        assert abs(a[2] - b[2]) < 1e-13 and abs(a[4] - b[4]) < 1e-13 * abs(b[4]) + 1e-16
    ref = full.psi.download()
    assert np.linalg.norm(rs.psi.download() - ref) < 1e-12 * np.linalg.norm(ref)
    assert not os.path.exists(os.path.join(rs.dir, "WF_0.wf"))
    w = Q.WaveFunction.load(ctx, os.path.join(rs.dir, "WF"), counter=21)
    assert np.array_equal(w.download(), rs.psi.download())
    assert os.path.exists(os.path.join(rs.dir, "WF_20.wf"))
    with pytest.raises(Q.QdxError):
        run("bad", 30, -1, os.path.join(first.dir, "WF_2"))
