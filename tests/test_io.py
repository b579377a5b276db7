"""QDng's binary file format (include/qdx_io.h) against QDng's own data files.

The fixtures under tests/golden/qdng_files are QDng regression-test inputs
copied unchanged (tests/golden/make_file_fixtures.py).  The raw values are
checked against numpy reading the same bytes and against the arrays of
tests/golden/qdng_kat.npz, which the oracle's regression tests consume; the
meta files QDng itself generated (GridGauss.meta, GridPot0_ef.meta) fix the
written format byte for byte.  CPU tests need no device; the `gpu` tests
load and save device wave functions and build Hamiltonians from files."""
import os

import numpy as np
import pytest

import cases
import qdng_amd as Q

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "qdng_files")


def fx(name):
    return os.path.join(FIX, name)


def uncommented(path):
    return "".join(l for l in open(path) if not l.lstrip().startswith("#"))


# ---------------------------------------------------------------------------
# meta files

def test_read_meta_of_qdng_files():
    assert Q.read_meta(fx("GridPot0.meta")) == {"CLASS": "GridPotential", "N0": "128", "dims": "1",
                                                "xmax0": "6.1104", "xmin0": "1.8897"}
    # a comment line holding '=' is skipped; unspaced "key=value" lines read
    assert Q.read_meta(fx("GridLaser_w0.176.meta")) == {"CLASS": "Laser", "Nt": "400", "dt": "1"}
    assert list(Q.read_meta(fx("GridGauss.meta"))) == ["CLASS", "N0", "dims", "xmax0", "xmin0"]


def test_meta_parse_rules(tmp_path):
    p = tmp_path / "x.meta"
    p.write_text("# note = ignored\n  a = 1 \nb[int] = 2\n=7\nno delimiter\nc=\t3\t\na = 4\nd = x = y\n")
    assert Q.read_meta(str(p)) == {"a": "4", "b": "2", "c": "3", "d": "x = y"}


def test_written_meta_matches_qdng_generated_files(tmp_path):
    # GridGauss.meta was written by QDng: keys in map order, "key = value",
    # doubles in 6 significant digits (its "# Generated file" line is a comment)
    out = tmp_path / "g.meta"
    Q.write_meta(str(out), {"xmin0": 1.8897, "dims": 1, "N0": 128, "xmax0": 6.1104, "CLASS": "WFGridCartesian"})
    assert out.read_text() == uncommented(fx("GridGauss.meta"))
    assert out.read_text() == uncommented(fx("GridPot0_ef.meta"))


def test_double_format_is_iostream_default(tmp_path):
    vals = [1.0 / 3.0, 1e-7, 123456789.0, -2.5, 10.0, 0.1 + 0.2, 6.02214076e23]
    out = tmp_path / "d.meta"
    Q.write_meta(str(out), {f"k{i}": v for i, v in enumerate(vals)})
    got = Q.read_meta(str(out))
    for i, v in enumerate(vals):
        assert got[f"k{i}"] == "%g" % v   # printf %g == ostream default (precision 6)


def test_meta_key_checks(tmp_path):
    for bad in ["", "#x", "a=b", "a[b", " a"]:
        with pytest.raises(Q.QdxError) as e:
            Q.write_meta(str(tmp_path / "k.meta"), {bad: "1"})
        assert e.value.code == -1
    with pytest.raises(Q.QdxError) as e:
        Q.read_meta(str(tmp_path / "absent.meta"))
    assert e.value.code == -7


# ---------------------------------------------------------------------------
# raw data files

def test_grid_operator_values():
    kat, _ = cases.kat_data()
    for name in ["GridPot0", "GridPot1", "GridDip0"]:
        N, xmin, xmax, v = Q.load_gridop(fx(name))
        assert (N, xmin, xmax) == ([128], [1.8897], [6.1104])
        assert np.array_equal(v, np.fromfile(fx(name + ".op"), dtype="<f8"))
    assert np.array_equal(Q.load_gridop(fx("GridPot0"))[3], kat["GridPot0"])


def test_potential_setzero_offset_scale_order():
    # OGridPotential.cpp:55-76: minimum to zero, then + offset, then * scale
    v = np.fromfile(fx("GridPot0.op"), dtype="<f8")
    got = Q.load_gridop(fx("GridPot0"), setzero=True, offset=0.25, scale=3.0)[3]
    assert np.array_equal(got, ((v - v.min()) + 0.25) * 3.0)
    assert not np.allclose(got, (v - v.min()) * 3.0 + 0.25)
    assert np.array_equal(Q.load_gridop(fx("GridPot0"), offset=-1.5)[3], v + -1.5)


def test_laser_file():
    kat, _ = cases.kat_data()
    dt, e = Q.load_laser(fx("GridLaser_w0.176"))
    assert dt == 1.0 and e.size == 400
    assert np.array_equal(e, kat["GridLaser_w0.176"])


def test_sequence_naming_and_meta_fallback():
    # GridPot0_ef_<n>.wf is a sequence with the meta file GridPot0_ef.meta
    kat, _ = cases.kat_data()
    for i in range(2):
        a = Q.read_file(fx("GridPot0_ef"), ".wf", 128, counter=i)
        assert np.array_equal(a, kat["GridPot0_ef_%d" % i])
        # the same file named without a sequence finds its meta file by
        # cutting the name at its last '_' (FileSingle.h:1217-1238)
        assert Q.file_meta(fx("GridPot0_ef_%d" % i)) == Q.read_meta(fx("GridPot0_ef.meta"))
        assert np.array_equal(Q.read_file(fx("GridPot0_ef_%d" % i), ".wf", 128), a)


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    psi = rng.standard_normal(96) + 1j * rng.standard_normal(96)
    base = str(tmp_path / "WF")
    Q.write_file(base, ".wf", psi, {"CLASS": "WFGridCartesian", "dims": 1, "N0": 96, "xmin0": -3.0,
                                    "xmax0": 3.0}, counter=5)
    assert os.path.exists(base + "_5.wf") and os.path.exists(base + ".meta")
    assert np.array_equal(Q.read_file(base, ".wf", 96, counter=5), psi)
    assert Q.file_meta(base)["N0"] == "96"
    # operator and laser writers read back by the readers (and by numpy)
    Q.save_gridop(str(tmp_path / "V"), [4, 8], [-1.0, 0.0], [1.0, 2.0], np.arange(32.0))
    N, xmin, xmax, v = Q.load_gridop(str(tmp_path / "V"))
    assert (N, xmin, xmax) == ([4, 8], [-1.0, 0.0], [1.0, 2.0]) and np.array_equal(v, np.arange(32.0))
    assert Q.read_meta(str(tmp_path / "V.meta"))["CLASS"] == "GridPotential"
    Q.save_laser(str(tmp_path / "L"), 0.5, np.linspace(0, 1, 7))
    dt, e = Q.load_laser(str(tmp_path / "L"))
    assert dt == 0.5 and np.array_equal(e, np.linspace(0, 1, 7))
    assert np.array_equal(np.fromfile(str(tmp_path / "L.op"), dtype="<f8"), e)


def test_propagation_meta(tmp_path):
    Q.write_propagation_meta(str(tmp_path) + "/", 1000, 0.5, "WF", 10)
    assert (tmp_path / "Propagation.meta").read_text() == \
        "CLASS = Propagation\nNt = 1000\nWFBaseName = WF\nWcycle = 10\ndt = 0.5\n"


def test_read_errors(tmp_path):
    # a longer file is read from its start, a shorter one is an error
    assert Q.read_file(fx("GridGauss"), ".wf", 64).size == 64
    with pytest.raises(Q.QdxError) as e:
        Q.read_file(fx("GridGauss"), ".wf", 129)
    assert e.value.code == -7
    # compressed payloads carry a tag in their last 8 bytes: a ZLIB tag
    # after data that is no zlib stream is an I/O error, BZIP is rejected
    for tag, code in ((b"ZLIBCOM\0", -7), (b"BZIPCOM\0", -2)):
        p = tmp_path / "c.wf"
        p.write_bytes(b"\x01" * 64 + tag)
        with pytest.raises(Q.QdxError) as e:
            Q.read_file(str(tmp_path / "c"), ".wf", 4)
        assert e.value.code == code
    with pytest.raises(Q.QdxError) as e:
        Q.file_meta(str(tmp_path / "none_3"))
    assert e.value.code == -7


# ---------------------------------------------------------------------------
# device objects from files

@pytest.mark.gpu
def test_wave_function_load_save(ctx, tmp_path):
    w = Q.WaveFunction.load(ctx, fx("GridGauss"))
    assert w.N == [128] and w.S == 1
    raw = np.fromfile(fx("GridGauss.wf"), dtype="<c16")
    assert np.array_equal(w.download(), raw)
    w.save(str(tmp_path / "G"))
    assert (tmp_path / "G.wf").read_bytes() == open(fx("GridGauss.wf"), "rb").read()
    assert (tmp_path / "G.meta").read_text() == uncommented(fx("GridGauss.meta"))


@pytest.mark.gpu
def test_multistate_wave_function_files(ctx, tmp_path):
    kat, _ = cases.kat_data()
    psi = np.concatenate([kat["GridPot0_ef_0"], kat["GridPot0_ef_1"]])
    w = Q.WaveFunction(ctx, [128], [1.8897], [6.1104], 2).upload(psi)
    base = str(tmp_path / "psi")
    w.save(base, counter=3)
    for f in ["psi-0_3.wf", "psi-1_3.wf", "psi-0.meta", "psi-1.meta", "psi.meta"]:
        assert (tmp_path / f).exists(), f
    assert Q.read_meta(base + ".meta") == {"CLASS": "WFMultistate", "states": "2"}
    w2 = Q.WaveFunction.load(ctx, base, counter=3)
    assert w2.S == 2 and w2.N == [128]
    assert np.array_equal(w2.download(), psi)
    with pytest.raises(Q.QdxError):
        Q.WaveFunction.load(ctx, base, counter=4)


@pytest.mark.gpu
def test_hamiltonian_terms_from_files(ctx):
    kat, _ = cases.kat_data()
    N, xmin, xmax = [128], [1.8897], [6.1104]
    v = np.fromfile(fx("GridPot0.op"), dtype="<f8")
    mu = np.fromfile(fx("GridDip0.op"), dtype="<f8")
    e = np.fromfile(fx("GridLaser_w0.176.op"), dtype="<f8")
    Ha = Q.Hamiltonian(ctx, N, xmin, xmax).add_kinetic(0, [1836.0])
    Ha.add_potential_file(0, 0, fx("GridPot0"), setzero=True, offset=0.1, scale=2.0)
    Ha.add_dipole_file(0, 0, fx("GridDip0"), fx("GridLaser_w0.176"), scale=0.5)
    Hb = Q.Hamiltonian(ctx, N, xmin, xmax).add_kinetic(0, [1836.0])
    Hb.add_potential(0, 0, ((v - v.min()) + 0.1) * 2.0)
    # QDng scales a GridDipole's array and its field (OGridPotential::Init,
    # OLaser::Clock): -scale^2 mu E
    Hb.add_dipole(0, 0, mu * 0.5, e, 0.5)
    src = Q.WaveFunction(ctx, N, xmin, xmax).upload(kat["GridPot0_ef_1"])
    da, db = Q.WaveFunction(ctx, N, xmin, xmax), Q.WaveFunction(ctx, N, xmin, xmax)
    for step in (0, 37):
        Ha.apply(da, src, step)
        Hb.apply(db, src, step)
        assert np.array_equal(da.download(), db.download())
    assert Ha.bounds() == Hb.bounds()
    # a file on another grid is rejected
    Hc = Q.Hamiltonian(ctx, [64], xmin, xmax)
    with pytest.raises(Q.QdxError) as ex:
        Hc.add_potential_file(0, 0, fx("GridPot0"))
    assert ex.value.code == -2
