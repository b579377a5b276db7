"""QDng's protobuf stream file format and ZLIB-compressed payloads.

The header messages are declared here, field by field, as
FileSingleHeader.proto and GridSystem.proto declare them, and encoded and
decoded with the protobuf library itself: the library's writer must produce
the same bytes as protobuf's serializer, and files that protobuf wrote must
load.  Payload layouts (FileSingle.h:950-1210, WFGridSystem.cpp:140-177,
OGridsystem.cpp:63-95, Laser.cpp:206-235, WFMultistate.cpp:279-293) are
built by hand in the tests.  Compressed wave functions carry QDng's k-space
reduction (WFGridSystem.cpp:90-137), checked against numpy's FFT.  CPU
tests cover lasers and grid operators; the `gpu` tests wave functions."""
import os
import struct
import zlib

import numpy as np
import pytest

import qdng_amd as Q


def _messages():
    from google.protobuf import descriptor_pb2, descriptor_pool, message_factory
    F = descriptor_pb2.FieldDescriptorProto
    fdp = descriptor_pb2.FileDescriptorProto(name="qdng_stream_test.proto", package="QDLIB", syntax="proto2")
    h = fdp.message_type.add(name="FileSingleHeader")
    e = h.enum_type.add(name="Compression")
    for n, v in (("UNCOMPRESSED", 0), ("ZLIB", 1), ("BZIP", 2)):
        e.value.add(name=n, number=v)
    for name, num, lab, typ, dflt in (("more_files_follow", 1, F.LABEL_REQUIRED, F.TYPE_BOOL, "false"),
                                      ("class", 2, F.LABEL_REQUIRED, F.TYPE_STRING, "NoClass"),
                                      ("name", 3, F.LABEL_OPTIONAL, F.TYPE_STRING, None),
                                      ("sequence_number", 4, F.LABEL_OPTIONAL, F.TYPE_UINT32, None),
                                      ("meta_data", 5, F.LABEL_OPTIONAL, F.TYPE_STRING, None),
                                      ("payload_size", 6, F.LABEL_REQUIRED, F.TYPE_UINT64, "0"),
                                      ("compression", 7, F.LABEL_REQUIRED, F.TYPE_ENUM, "UNCOMPRESSED"),
                                      ("sub_header", 10, F.LABEL_OPTIONAL, F.TYPE_BYTES, None)):
        f = h.field.add(name=name, number=num, label=lab, type=typ)
        if typ == F.TYPE_ENUM:
            f.type_name = ".QDLIB.FileSingleHeader.Compression"
        if dflt is not None:
            f.default_value = dflt
    g = fdp.message_type.add(name="GridSystemHeader")
    d = g.nested_type.add(name="dim_description_t")
    for name, num, typ in (("size", 1, F.TYPE_UINT32), ("xmin", 2, F.TYPE_DOUBLE), ("xmax", 3, F.TYPE_DOUBLE)):
        d.field.add(name=name, number=num, label=F.LABEL_REQUIRED, type=typ, default_value="0")
    g.field.add(name="dims", number=1, label=F.LABEL_REQUIRED, type=F.TYPE_UINT32, default_value="0")
    g.field.add(name="dim", number=2, label=F.LABEL_REPEATED, type=F.TYPE_MESSAGE,
                type_name=".QDLIB.GridSystemHeader.dim_description_t")
    g.field.add(name="data_size", number=3, label=F.LABEL_OPTIONAL, type=F.TYPE_UINT64, default_value="0")
    pool = descriptor_pool.DescriptorPool()
    pool.Add(fdp)
    return (message_factory.GetMessageClass(pool.FindMessageTypeByName("QDLIB.FileSingleHeader")),
            message_factory.GetMessageClass(pool.FindMessageTypeByName("QDLIB.GridSystemHeader")))


Header, GridHeader = _messages()
INTRO = b"QDngPrtoBuf\0" + struct.pack("<I", 1)     # PbufHeader (FileSingle.h:141-147)


def read_entries(path):
    """(header message, its bytes, payload) per entry of a stream file."""
    b = open(path, "rb").read()
    assert b[:16] == INTRO
    at, out = 16, []
    while True:
        (n,) = struct.unpack_from("<I", b, at)
        hb = b[at + 4:at + 4 + n]
        h = Header()
        h.ParseFromString(hb)                 # required fields must be present
        at += 4 + n
        pl = b[at:at + h.payload_size]
        at += h.payload_size
        out.append((h, hb, pl))
        if not h.more_files_follow:
            break
    assert at == len(b)
    return out


def entry(cls, name, meta, payload, more=False, comp=0):
    h = Header()
    h.more_files_follow = more
    setattr(h, "class", cls)
    h.name = name
    h.meta_data = meta
    pl = zlib.compress(payload, 4) if comp == 1 else payload
    h.payload_size = len(pl)
    h.compression = comp
    hb = h.SerializeToString()
    return struct.pack("<I", len(hb)) + hb + pl


def grid_payload(N, lo, hi, values):
    g = GridHeader()
    g.dims = len(N)
    for n, a, b in zip(N, lo, hi):
        d = g.dim.add()
        d.size, d.xmin, d.xmax = n, a, b
    v = np.ascontiguousarray(values).view(np.float64).tobytes()
    g.data_size = len(v)
    gb = g.SerializeToString()
    return struct.pack("<I", len(gb)) + gb + v


@pytest.fixture
def stream_format():
    old = Q.get_file_format()
    Q.set_file_format("stream")
    yield
    Q.set_file_format(old[0], old[1], old[2])


def test_laser_stream_bytes_match_protobuf(tmp_path, stream_format):
    field = np.sin(np.arange(37) * 0.3) * 0.01
    name = str(tmp_path / "laser")
    Q.save_laser(name, 0.5, field)
    path = name + ".op"
    assert Q.file_format(path) == "stream"
    ((h, hb, pl),) = read_entries(path)
    assert getattr(h, "class") == "Laser" and h.name == path and not h.more_files_follow
    assert h.meta_data == "Nt = 37\ndt = 0.5\n" and h.compression == 0
    assert hb == h.SerializeToString()        # protobuf's own encoding of the same message
    n, dt = struct.unpack_from("<Qd", pl, 0)
    assert (n, dt) == (37, 0.5)
    assert np.array_equal(np.frombuffer(pl[16:], np.float64), field)
    assert Q.load_laser(name)[0] == 0.5 and np.array_equal(Q.load_laser(name)[1], field)


def test_gridop_stream_round_trip_and_protobuf_decode(tmp_path, stream_format):
    N, lo, hi = [12, 5], [-1.0, 0.5], [2.0, 3.25]
    v = np.arange(60.0) ** 1.5
    name = str(tmp_path / "pot")
    Q.save_gridop(name, N, lo, hi, v)
    ((h, hb, pl),) = read_entries(name + ".op")
    assert getattr(h, "class") == "OGridPotential" and hb == h.SerializeToString()
    (gl,) = struct.unpack_from("<I", pl, 0)
    g = GridHeader()
    g.ParseFromString(pl[4:4 + gl])
    assert g.dims == 2 and [d.size for d in g.dim] == N and [d.xmin for d in g.dim] == lo
    assert [d.xmax for d in g.dim] == hi and g.data_size == 60 * 8
    assert pl[4:4 + gl] == g.SerializeToString()
    assert np.array_equal(np.frombuffer(pl[4 + gl:], np.float64), v)
    N2, lo2, hi2, v2 = Q.load_gridop(name, setzero=True, offset=1.0, scale=2.0)
    assert (N2, lo2, hi2) == (N, lo, hi)
    assert np.allclose(v2, 2.0 * (v - v.min() + 1.0), rtol=0, atol=0)


@pytest.mark.parametrize("comp", [0, 1])
def test_files_written_by_protobuf_load(tmp_path, comp):
    # a laser and a potential as protobuf writes them (ZLIB payloads too)
    field = np.linspace(-1, 1, 11)
    pl = struct.pack("<Qd", 11, 2.0) + field.tobytes()
    (tmp_path / "las.op").write_bytes(INTRO + entry("Laser", "las.op", "Nt = 11\ndt = 2\n", pl, comp=comp))
    dt, got = Q.load_laser(str(tmp_path / "las"))
    assert dt == 2.0 and np.array_equal(got, field)
    v = np.cos(np.arange(16.0))
    (tmp_path / "v.op").write_bytes(INTRO + entry("OGridPotential", "v.op", "", grid_payload([16], [0.0], [3.0], v),
                                                  comp=comp))
    N, lo, hi, got = Q.load_gridop(str(tmp_path / "v"))
    assert N == [16] and lo == [0.0] and hi == [3.0] and np.array_equal(got, v)


def test_raw_zlib_payload_and_rejections(tmp_path):
    v = np.arange(20.0)
    z = zlib.compress(v.tobytes(), 4) + b"ZLIBCOM\0"   # FileSingle.h:1336-1350
    (tmp_path / "a.op").write_bytes(z)
    out = np.zeros(20)
    Q.lib().qdx_file_read(str(tmp_path / "a").encode(), b".op", -1, out.ctypes.data_as(Q.C.c_void_p), 160)
    assert np.array_equal(out, v)
    (tmp_path / "b.op").write_bytes(b"\0" * 40 + b"BZIPCOM\0")
    with pytest.raises(Q.QdxError) as e:
        Q._chk(Q.lib().qdx_file_read(str(tmp_path / "b").encode(), b".op", -1, out.ctypes.data_as(Q.C.c_void_p),
                                     160))
    assert e.value.code == -2
    (tmp_path / "c.op").write_bytes(INTRO + b"\x05\0\0\0\xff\xff")     # truncated header
    with pytest.raises(Q.QdxError):
        Q.load_laser(str(tmp_path / "c"))
    with pytest.raises(Q.QdxError):
        Q.set_file_format("hdf5")


# ---------------------------------------------------------------------------
# wave functions (device)

def _restore_reduce(psi, shape_np, tol):
    """numpy: WFGridSystem::Reduce then Restore (forward transform over every
    dimension, components below tol * max|psi~| zeroed, / size, backward)."""
    a = np.fft.fftn(psi.reshape(shape_np))
    thr = np.abs(a).max() * tol
    re = np.where(np.abs(a.real) < thr, 0.0, a.real / a.size)
    im = np.where(np.abs(a.imag) < thr, 0.0, a.imag / a.size)
    return np.fft.ifftn(re + 1j * im).ravel() * a.size


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,compress", [("stream", False), ("stream", True), ("binary", True)])
@pytest.mark.parametrize("S", [1, 2])
def test_wave_function_formats(ctx, tmp_path, fmt, compress, S):
    print(f"[gpu case] wf file {fmt} compress={compress} S={S}", flush=True)
    N, lo, hi = [24, 16], [-3.0, -1.0], [3.0, 2.0]
    rng = np.random.default_rng(5)
    x = np.linspace(-3, 3, 24)[None, :]
    psi = np.exp(-x ** 2) * (1 + 0.3 * rng.standard_normal((16, 24))) + 0j
    data = np.concatenate([psi.ravel() * (k + 1) for k in range(S)])
    w = Q.WaveFunction(ctx, N, lo, hi, S)
    w.upload(data)
    old = Q.get_file_format()
    try:
        Q.set_file_format(fmt, compress, 1e-3)
        w.save(str(tmp_path / "WF"), counter=4)
    finally:
        Q.set_file_format(*old)
    path = str(tmp_path / "WF_4.wf") if (fmt == "stream" or S == 1) else str(tmp_path / "WF-0_4.wf")
    assert Q.file_format(path) == fmt
    got = Q.WaveFunction.load(ctx, str(tmp_path / "WF"), counter=4).download()
    if compress:
        want = np.concatenate([_restore_reduce(data[k * 384:(k + 1) * 384], (16, 24), 1e-3) for k in range(S)])
        assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))
        assert 1e-6 < np.max(np.abs(got - data)) / np.max(np.abs(data)) < 1e-2   # the reduction is lossy
    else:
        assert np.array_equal(got, data)
    if fmt == "stream":
        es = read_entries(path)
        assert len(es) == S + (1 if S > 1 else 0)
        if S > 1:
            h, hb, pl = es[0]
            assert getattr(h, "class") == "WFMultistate" and h.more_files_follow
            raw = zlib.decompress(pl) if compress else pl
            assert struct.unpack("<I", raw) == (S,)
        for k, (h, hb, pl) in enumerate(es[-S:]):
            assert getattr(h, "class") == "WFGridCartesian" and hb == h.SerializeToString()
            assert h.more_files_follow == (k < S - 1) and h.compression == int(compress)
            assert h.meta_data == "N0 = 24\nN1 = 16\ndims = 2\nxmax0 = 3\nxmax1 = 2\nxmin0 = -3\nxmin1 = -1\n"


@pytest.mark.gpu
def test_multistate_stream_written_by_protobuf_loads(ctx, tmp_path):
    N, lo, hi = [16], [0.0], [1.0]
    a = np.exp(1j * np.arange(16)) * np.arange(16)
    b = np.conj(a) * 0.5
    f = INTRO + entry("WFMultistate", "m.wf", "states = 2\n", struct.pack("<I", 2), more=True)
    f += entry("WFGridCartesian", "m.wf", "", grid_payload(N, lo, hi, a), more=True)
    f += entry("WFGridCartesian", "m.wf", "", grid_payload(N, lo, hi, b))
    (tmp_path / "m.wf").write_bytes(f)
    w = Q.WaveFunction.load(ctx, str(tmp_path / "m"))
    assert w.S == 2 and np.array_equal(w.download(), np.concatenate([a, b]))
