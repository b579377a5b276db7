// QDng's file formats on the host: key/value meta files (ParamContainer),
// raw data files with sequence naming (FileSingle, binary format), the
// protobuf stream format (FileSingle, stream format: FileSingleHeader and
// GridSystemHeader messages in protobuf's wire encoding, written here
// without the protobuf library), ZLIB-compressed payloads with QDng's
// k-space reduction of wave functions, operator and laser files,
// Propagation.meta, and wave functions / Hamiltonian terms loaded from them.
// The C interface is include/qdx_io.h.
#include <zlib.h>

#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "../../include/qdx_io.h"
#include "qdx_host_internal.hpp"

struct qdx_meta {
    std::map<std::string, std::string> kv;   // ordered as ParamContainer's std::map
};

namespace {

constexpr const char *kMetaSuffix = ".meta";   // FileSingle.h:51
constexpr const char *kWfSuffix = ".wf";       // FileSingle.h:45
constexpr const char *kOpSuffix = ".op";       // FileSingle.h:46
constexpr int kTagLen = 8;                     // compression tag "ZLIBCOM" / "BZIPCOM" + NUL

// helpers.cpp:9-30: strip blanks and tabs at both ends
std::string trim(const std::string &s) {
    const size_t a = s.find_first_not_of(" \t");
    if (a == std::string::npos) return std::string();
    const size_t b = s.find_last_not_of(" \t");
    return s.substr(a, b - a + 1);
}

void parse_meta(std::istream &in, qdx_meta &m) {
    std::string line;
    while (std::getline(in, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq));
        const std::string val = trim(line.substr(eq + 1));
        const size_t br = key.find('[');   // "key[type] = value": the type is dropped
        if (br != std::string::npos) key = trim(key.substr(0, br));
        if (key.empty() || key[0] == '#') continue;
        m.kv[key] = val;
    }
}

bool read_meta_file(const std::string &path, qdx_meta &m) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    parse_meta(f, m);
    return true;
}

void write_meta_file(const qdx_meta &m, const std::string &path) {
    std::ofstream f(path, std::ios::out | std::ios::trunc);
    if (!f.is_open()) fail(QDX_ERR_IO, "Can not write meta file: " + path);
    for (const auto &kv : m.kv) f << kv.first << " = " << kv.second << "\n";
    f.close();
    if (f.fail()) fail(QDX_ERR_IO, "Can not write meta file: " + path);
}

// a double as an iostream with default flags prints it (ParamContainer.cpp:584-592)
std::string fmt_double(double v) {
    std::ostringstream ss;
    ss << v;
    return ss.str();
}

const std::string &meta_need(const qdx_meta &m, const std::string &key, const std::string &what) {
    auto it = m.kv.find(key);
    if (it == m.kv.end()) fail(QDX_ERR_PARAM, what + ": key '" + key + "' missing");
    return it->second;
}

long meta_int(const qdx_meta &m, const std::string &key, const std::string &what) {
    const std::string &s = meta_need(m, key, what);
    char *end = nullptr;
    errno = 0;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || errno) fail(QDX_ERR_PARAM, what + ": '" + key + "' is not an integer: " + s);
    return v;
}

double meta_double(const qdx_meta &m, const std::string &key, const std::string &what) {
    const std::string &s = meta_need(m, key, what);
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) fail(QDX_ERR_PARAM, what + ": '" + key + "' is not a number: " + s);
    return v;
}

std::string need_str(const char *s, const char *what) {
    if (!s) fail(QDX_ERR_PARAM, std::string(what) + " is NULL");
    return s;
}

// FileSingle::BuildDataFileName (FileSingle.h:1398-1410)
std::string data_name(const std::string &name, const std::string &suffix, int counter) {
    if (counter < 0) return name + suffix;
    return name + "_" + std::to_string(counter) + suffix;
}

// FileSingle::ReadMetaFromRaw (FileSingle.h:1217-1238)
qdx_meta file_meta(const std::string &name) {
    qdx_meta m;
    if (read_meta_file(name + kMetaSuffix, m)) return m;
    const size_t us = name.rfind('_');
    if (us != std::string::npos && read_meta_file(name.substr(0, us) + kMetaSuffix, m)) return m;
    fail(QDX_ERR_IO, "Can not read meta file " + name + kMetaSuffix + (us != std::string::npos ? " or " + name.substr(0, us) + kMetaSuffix : ""));
}

std::string zlib_inflate(const char *p, size_t n, size_t expect);

// FileSingle::DetectCompressionInPayload + ReadDataFromRaw (FileSingle.h:
// 1245-1287, 769-801): exactly `bytes` bytes, inflated when the file ends in
// the ZLIB tag; returns whether it was compressed
bool read_raw(const std::string &path, void *dst, size_t bytes) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) fail(QDX_ERR_IO, "Can not open binary file for reading: " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff len = f.tellg();
    if (len < kTagLen) fail(QDX_ERR_IO, "Binary file truncated: " + path);
    char tag[kTagLen];
    f.seekg(len - kTagLen);
    f.read(tag, kTagLen);
    if (std::strncmp(tag, "BZIPCOM", kTagLen) == 0)
        fail(QDX_ERR_INCOMPAT, "BZIP-compressed binary file: not supported (no libbz2): " + path);
    if (std::strncmp(tag, "ZLIBCOM", kTagLen) == 0) {
        std::string all((size_t)len, '\0');
        f.seekg(0);
        f.read(&all[0], len);
        const std::string out = zlib_inflate(all.data(), all.size(), bytes);
        if (out.size() != bytes) fail(QDX_ERR_IO, "zlib decompression failed: " + path);
        std::memcpy(dst, out.data(), bytes);
        return true;
    }
    f.seekg(0);
    f.read(static_cast<char *>(dst), (std::streamsize)bytes);
    if ((size_t)f.gcount() != bytes) fail(QDX_ERR_IO, "Can not read binary file: " + path);
    return false;
}

void write_raw(const std::string &path, const void *src, size_t bytes) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.is_open()) fail(QDX_ERR_IO, "Can not open binary file for writing: " + path);
    f.write(static_cast<const char *>(src), (std::streamsize)bytes);
    f.close();
    if (f.fail()) fail(QDX_ERR_IO, "Can not write binary file: " + path);
}

// grid keys of GridSystemHeaderPm (GridSystem.cpp:17-70)
qdx_grid grid_from_meta(const qdx_meta &m, const std::string &what) {
    qdx_grid g{};
    g.ndims = (int32_t)meta_int(m, "dims", what);
    if (g.ndims < 1 || g.ndims > 4) fail(QDX_ERR_PARAM, what + ": dims out of range");
    for (int d = 0; d < g.ndims; ++d) {
        const std::string c = std::to_string(d);
        g.N[d] = (int32_t)meta_int(m, "N" + c, what);
        if (g.N[d] < 1) fail(QDX_ERR_PARAM, what + ": Zero points grid defined");
        g.xmin[d] = meta_double(m, "xmin" + c, what);
        g.xmax[d] = meta_double(m, "xmax" + c, what);
        if (g.xmax[d] - g.xmin[d] <= 0) fail(QDX_ERR_PARAM, what + ": Zero length grid defined");
    }
    return g;
}

void grid_to_meta(const qdx_grid &g, qdx_meta &m) {
    m.kv["dims"] = std::to_string(g.ndims);
    for (int d = 0; d < g.ndims; ++d) {
        const std::string c = std::to_string(d);
        m.kv["xmin" + c] = fmt_double(g.xmin[d]);
        m.kv["xmax" + c] = fmt_double(g.xmax[d]);
        m.kv["N" + c] = std::to_string(g.N[d]);
    }
}

long grid_points(const qdx_grid &g) {
    long n = 1;
    for (int d = 0; d < g.ndims; ++d) n *= g.N[d];
    return n;
}

// the grid of a file against an object's grid: same dimensions and sizes,
// end points equal to the digits a meta file carries
void check_same_grid(const qdx_grid &f, const Grid &g, const std::string &what) {
    bool same = f.ndims == g.nd;
    for (int d = 0; same && d < g.nd; ++d) {
        same = f.N[d] == g.N[d];
        const double tol = 1e-9 * std::max(1.0, std::max(std::fabs(g.xmin[d]), std::fabs(g.xmax[d])));
        same = same && std::fabs(f.xmin[d] - g.xmin[d]) <= tol && std::fabs(f.xmax[d] - g.xmax[d]) <= tol;
    }
    if (!same) fail(QDX_ERR_INCOMPAT, what + ": the file's grid differs from the operator's");
}

qdx_grid to_qdx_grid(const Grid &g) {
    qdx_grid q{};
    q.ndims = g.nd;
    for (int d = 0; d < g.nd; ++d) {
        q.N[d] = g.N[d];
        q.xmin[d] = g.xmin[d];
        q.xmax[d] = g.xmax[d];
    }
    return q;
}

// ---------------------------------------------------------------------------
// output settings (QDng's global parameters wf_format, compress, comptol;
// FileWF.cpp:17-56)
struct IoFormat {
    int wf_format = QDX_FORMAT_BINARY;
    int compress = 0;
    double comptol = 1e-6;    // FILEWF_COMPRESSION_TOLERANCE (FileWF.h:12)
};
IoFormat g_io;

// ---------------------------------------------------------------------------
// protobuf wire encoding of the two header messages (FileSingleHeader.proto,
// GridSystem.proto): varint keys (field << 3 | wire type), varints, 64-bit
// little-endian doubles, length-delimited strings and sub-messages; fields
// in field-number order, as protobuf's serializer emits them
constexpr int kVarint = 0, kFixed64 = 1, kLen = 2, kFixed32 = 5;

struct PbOut {
    std::string b;
    void varint(uint64_t v) {
        while (v >= 0x80) {
            b.push_back(char((v & 0x7f) | 0x80));
            v >>= 7;
        }
        b.push_back(char(v));
    }
    void key(int field, int wt) { varint((uint64_t(field) << 3) | uint64_t(wt)); }
    void u64(int field, uint64_t v) {
        key(field, kVarint);
        varint(v);
    }
    void f64(int field, double v) {
        key(field, kFixed64);
        char c[8];
        std::memcpy(c, &v, 8);
        b.append(c, 8);
    }
    void bytes(int field, const std::string &v) {
        key(field, kLen);
        varint(v.size());
        b += v;
    }
};

struct PbIn {
    const unsigned char *p, *e;
    explicit PbIn(const std::string &s)
        : p(reinterpret_cast<const unsigned char *>(s.data())), e(p + s.size()) {}
    [[noreturn]] static void bad() { fail(QDX_ERR_IO, "Error reading meta data from stream"); }
    uint64_t varint() {
        uint64_t v = 0;
        for (int sh = 0; sh < 64; sh += 7) {
            if (p >= e) bad();
            const unsigned c = *p++;
            v |= uint64_t(c & 0x7f) << sh;
            if (!(c & 0x80)) return v;
        }
        bad();
    }
    bool next(int &field, int &wt) {
        if (p >= e) return false;
        const uint64_t k = varint();
        field = int(k >> 3);
        wt = int(k & 7);
        return true;
    }
    double f64() {
        if (e - p < 8) bad();
        double v;
        std::memcpy(&v, p, 8);
        p += 8;
        return v;
    }
    std::string bytes() {
        const uint64_t n = varint();
        if (uint64_t(e - p) < n) bad();
        std::string v(reinterpret_cast<const char *>(p), size_t(n));
        p += n;
        return v;
    }
    void skip(int wt) {
        if (wt == kVarint) varint();
        else if (wt == kFixed64) f64();
        else if (wt == kLen) bytes();
        else if (wt == kFixed32 && e - p >= 4) p += 4;
        else bad();
    }
};

// FileSingleHeader (FileSingleHeader.proto): more_files_follow = 1,
// class = 2, name = 3, sequence_number = 4, meta_data = 5,
// payload_size = 6, compression = 7, sub_header = 10
struct StreamHeader {
    bool more = false;
    std::string cls = "NoClass", name, meta;
    uint64_t payload = 0;
    int comp = 0;   // 0 UNCOMPRESSED, 1 ZLIB, 2 BZIP
};

std::string encode_header(const StreamHeader &h) {   // the fields WriteMetaToStream sets
    PbOut o;
    o.u64(1, h.more ? 1 : 0);
    o.bytes(2, h.cls);
    o.bytes(3, h.name);
    o.bytes(5, h.meta);
    o.u64(6, h.payload);
    o.u64(7, uint64_t(h.comp));
    return o.b;
}

StreamHeader decode_header(const std::string &s) {
    StreamHeader h;
    PbIn in(s);
    int f, wt;
    while (in.next(f, wt)) {
        if (f == 1 && wt == kVarint) h.more = in.varint() != 0;
        else if (f == 2 && wt == kLen) h.cls = in.bytes();
        else if (f == 3 && wt == kLen) h.name = in.bytes();
        else if (f == 5 && wt == kLen) h.meta = in.bytes();
        else if (f == 6 && wt == kVarint) h.payload = in.varint();
        else if (f == 7 && wt == kVarint) h.comp = int(in.varint());
        else in.skip(wt);
    }
    return h;
}

// GridSystemHeader (GridSystem.proto): dims = 1, dim = 2 (size = 1,
// xmin = 2, xmax = 3), data_size = 3
std::string encode_grid(const qdx_grid &g, uint64_t data_size) {
    PbOut o;
    o.u64(1, uint64_t(g.ndims));
    for (int d = 0; d < g.ndims; ++d) {
        PbOut m;
        m.u64(1, uint64_t(g.N[d]));
        m.f64(2, g.xmin[d]);
        m.f64(3, g.xmax[d]);
        o.bytes(2, m.b);
    }
    o.u64(3, data_size);
    return o.b;
}

qdx_grid decode_grid(const std::string &s) {
    qdx_grid g{};
    PbIn in(s);
    int f, wt, nd = 0;
    while (in.next(f, wt)) {
        if (f == 1 && wt == kVarint) {
            g.ndims = int32_t(in.varint());
        } else if (f == 2 && wt == kLen) {
            if (nd >= 4) fail(QDX_ERR_PARAM, "stream grid: more than 4 dimensions");
            const std::string sub = in.bytes();
            PbIn m(sub);
            int f2, w2;
            while (m.next(f2, w2)) {
                if (f2 == 1 && w2 == kVarint) g.N[nd] = int32_t(m.varint());
                else if (f2 == 2 && w2 == kFixed64) g.xmin[nd] = m.f64();
                else if (f2 == 3 && w2 == kFixed64) g.xmax[nd] = m.f64();
                else m.skip(w2);
            }
            ++nd;
        } else {
            in.skip(wt);
        }
    }
    if (g.ndims < 1 || g.ndims > 4 || nd != g.ndims) fail(QDX_ERR_PARAM, "stream grid: invalid dimensions");
    for (int d = 0; d < g.ndims; ++d)
        if (g.N[d] < 1 || !(g.xmax[d] - g.xmin[d] > 0)) fail(QDX_ERR_PARAM, "stream grid: invalid dimension");
    return g;
}

// the intro of a stream (PbufHeader, FileSingle.h:53-54, 141-147): the
// zero-terminated "QDngPrtoBuf" and the 32-bit version 1
constexpr char kPbufMagic[] = "QDngPrtoBuf";
constexpr uint32_t kPbufVersion = 1;

std::string stream_intro() {
    std::string s(kPbufMagic, sizeof(kPbufMagic));
    char v[4];
    std::memcpy(v, &kPbufVersion, 4);
    s.append(v, 4);
    return s;
}

// FileSingle::DetectFormat (FileSingle.h:1110-1140): a file that opens with
// the intro is a stream, anything else the binary format
bool is_stream(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return false;
    char buf[sizeof(kPbufMagic) + 4];
    f.read(buf, sizeof(buf));
    if (f.gcount() != (std::streamsize)sizeof(buf)) return false;
    uint32_t v;
    std::memcpy(&v, buf + sizeof(kPbufMagic), 4);
    return std::string(buf) == kPbufMagic && v == kPbufVersion;
}

struct StreamEntry {
    StreamHeader h;
    std::string payload;   // inflated when it was compressed
};

// zlib (FileSingle.h:595-670): level 4, zlib format
std::string zlib_deflate(const std::string &in) {
    uLongf n = compressBound(in.size());
    std::string out(n, '\0');
    if (compress2(reinterpret_cast<Bytef *>(&out[0]), &n, reinterpret_cast<const Bytef *>(in.data()), in.size(), 4) !=
        Z_OK)
        fail(QDX_ERR_IO, "zlib deflate failed");
    out.resize(n);
    return out;
}

// inflate a zlib stream that may be followed by other bytes (the raw
// format's tag); `expect` bytes when known (> 0)
std::string zlib_inflate(const char *p, size_t n, size_t expect) {
    z_stream z{};
    if (inflateInit(&z) != Z_OK) fail(QDX_ERR_IO, "z-lib init failed");
    std::string out;
    out.resize(expect > 0 ? expect : 4 * n + 1024);
    z.next_in = reinterpret_cast<Bytef *>(const_cast<char *>(p));
    z.avail_in = uInt(n);
    int rc = Z_OK;
    size_t done = 0;
    while (true) {
        z.next_out = reinterpret_cast<Bytef *>(&out[done]);
        z.avail_out = uInt(out.size() - done);
        rc = inflate(&z, Z_NO_FLUSH);
        done = out.size() - z.avail_out;
        if (rc == Z_STREAM_END) break;
        if (rc != Z_OK && rc != Z_BUF_ERROR) break;
        if (z.avail_out == 0) {
            if (expect > 0) break;
            out.resize(out.size() * 2);
        } else if (z.avail_in == 0) {
            break;
        }
    }
    inflateEnd(&z);
    if (rc != Z_STREAM_END) fail(QDX_ERR_IO, "zlib decompression failed");
    out.resize(done);
    return out;
}

std::vector<StreamEntry> read_stream(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) fail(QDX_ERR_IO, "Can not open binary stream file for reading: " + path);
    std::string intro(stream_intro().size(), '\0');
    f.read(&intro[0], intro.size());
    if (intro != stream_intro()) fail(QDX_ERR_IO, "Malformated stream. No stream magic found.");
    std::vector<StreamEntry> out;
    while (true) {
        uint32_t mlen = 0;
        f.read(reinterpret_cast<char *>(&mlen), 4);
        if (f.gcount() != 4) fail(QDX_ERR_IO, "Error reading meta data from stream: " + path);
        std::string hb(mlen, '\0');
        f.read(&hb[0], mlen);
        if ((uint32_t)f.gcount() != mlen) fail(QDX_ERR_IO, "Error reading meta data from stream: " + path);
        StreamEntry e;
        e.h = decode_header(hb);
        e.payload.resize(e.h.payload);
        f.read(&e.payload[0], (std::streamsize)e.h.payload);
        if ((uint64_t)f.gcount() != e.h.payload) fail(QDX_ERR_IO, "Reading from stream failed: " + path);
        if (e.h.comp == 2) fail(QDX_ERR_INCOMPAT, "BZIP-compressed stream: not supported (no libbz2): " + path);
        if (e.h.comp == 1) e.payload = zlib_inflate(e.payload.data(), e.payload.size(), 0);
        const bool more = e.h.more;
        out.push_back(std::move(e));
        if (!more) break;
    }
    return out;
}

// ParamContainer::Parse of a header's meta_data
qdx_meta meta_of(const StreamHeader &h) {
    qdx_meta m;
    std::istringstream in(h.meta);
    parse_meta(in, m);
    return m;
}

std::string meta_text(const qdx_meta &m) {   // ParamContainer::Write, no types
    std::ostringstream ss;
    for (const auto &kv : m.kv) ss << kv.first << " = " << kv.second << "\n";
    return ss.str();
}

uint32_t get_u32(const std::string &s, size_t &at) {
    if (s.size() < at + 4) fail(QDX_ERR_IO, "Can't read WF from stream");
    uint32_t v;
    std::memcpy(&v, s.data() + at, 4);
    at += 4;
    return v;
}

// WFGridSystem / OGridSystem::DeSerialize (WFGridSystem.cpp:158-177,
// OGridsystem.cpp:78-95): a length-prefixed GridSystemHeader, then the
// values (ncomp doubles per point)
std::vector<double> grid_payload(const std::string &pl, int ncomp, qdx_grid &g) {
    size_t at = 0;
    const uint32_t hl = get_u32(pl, at);
    if (pl.size() < at + hl) fail(QDX_ERR_IO, "Can't read WF from stream");
    g = decode_grid(pl.substr(at, hl));
    at += hl;
    std::vector<double> v(size_t(grid_points(g)) * ncomp);
    if (pl.size() < at + v.size() * 8) fail(QDX_ERR_IO, "stream payload truncated");
    std::memcpy(v.data(), pl.data() + at, v.size() * 8);
    return v;
}

std::string grid_payload_out(const qdx_grid &g, const double *v, int ncomp) {
    const uint64_t bytes = uint64_t(grid_points(g)) * ncomp * 8;
    const std::string hdr = encode_grid(g, bytes);
    std::string pl;
    const uint32_t hl = uint32_t(hdr.size());
    pl.append(reinterpret_cast<const char *>(&hl), 4);
    pl += hdr;
    pl.append(reinterpret_cast<const char *>(v), bytes);
    return pl;
}

// one stream entry: the header (its length first), then the payload
// (WriteMetaToStream + WriteDataToStream, FileSingle.h:821-860, 983-987)
void put_entry(std::string &out, StreamHeader h, const std::string &payload) {
    std::string pl = h.comp == 1 ? zlib_deflate(payload) : payload;
    h.payload = pl.size();
    const std::string hb = encode_header(h);
    const uint32_t hl = uint32_t(hb.size());
    out.append(reinterpret_cast<const char *>(&hl), 4);
    out += hb;
    out += pl;
}

// ---------------------------------------------------------------------------
// QDng's lossy k-space reduction of compressed wave functions
// (WFGridSystem::Reduce / Restore, WFGridSystem.cpp:90-137): the forward
// transform (e^{-i}, unnormalised, every dimension), components below
// tolerance * max |psi~| set to zero, the rest divided by the point count;
// Restore is the unnormalised backward transform.  Host transforms: mixed
// radix (2, 3, 5, 7) recursion with a direct sum for other prime factors.
using cplx = std::complex<double>;

void dft_rec(const cplx *in, cplx *out, long n, long stride, int sign, std::vector<cplx> &scratch) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    long r = 0;
    for (long q : {4L, 2L, 3L, 5L, 7L})
        if (n % q == 0) {
            r = q;
            break;
        }
    if (r == 0) {   // a prime factor outside the radix set: direct sum
        for (long k = 0; k < n; ++k) {
            cplx s = 0;
            for (long j = 0; j < n; ++j) s += in[j * stride] * std::polar(1.0, sign * 2 * M_PI * double((j * k) % n) / double(n));
            out[k] = s;
        }
        return;
    }
    const long m = n / r;
    for (long q = 0; q < r; ++q) dft_rec(in + q * stride, out + q * m, m, stride * r, sign, scratch);
    std::vector<cplx> t(r);
    for (long k = 0; k < m; ++k) {
        for (long q = 0; q < r; ++q) t[q] = out[q * m + k] * std::polar(1.0, sign * 2 * M_PI * double(q * k) / double(n));
        for (long q = 0; q < r; ++q) {
            cplx s = 0;
            for (long j = 0; j < r; ++j) s += t[j] * std::polar(1.0, sign * 2 * M_PI * double((j * q) % r) / double(r));
            scratch[q * m + k] = s;
        }
    }
    for (long k = 0; k < n; ++k) out[k] = scratch[k];
}

void fft_grid(std::vector<cplx> &a, const qdx_grid &g, int sign) {
    long stride = 1;
    for (int d = 0; d < g.ndims; ++d) {
        const long L = g.N[d], total = (long)a.size();
        std::vector<cplx> line(L), res(L), scratch(L);
        for (long base = 0; base < total; ++base) {
            if ((base / stride) % L != 0) continue;   // first point of a line along d
            for (long i = 0; i < L; ++i) line[i] = a[base + i * stride];
            dft_rec(line.data(), res.data(), L, 1, sign, scratch);
            for (long i = 0; i < L; ++i) a[base + i * stride] = res[i];
        }
        stride *= L;
    }
}

std::vector<double> wf_reduce(const double *psi, const qdx_grid &g, double tol) {
    const long n = grid_points(g);
    std::vector<cplx> a(n);
    for (long i = 0; i < n; ++i) a[i] = cplx(psi[2 * i], psi[2 * i + 1]);
    fft_grid(a, g, -1);
    double mx = 0;
    for (const cplx &z : a) mx = std::max(mx, std::abs(z));
    const double thr = mx * tol;
    std::vector<double> out(2 * n);
    for (long i = 0; i < n; ++i) {
        out[2 * i] = std::fabs(a[i].real()) < thr ? 0.0 : a[i].real() / double(n);
        out[2 * i + 1] = std::fabs(a[i].imag()) < thr ? 0.0 : a[i].imag() / double(n);
    }
    return out;
}

void wf_restore(std::vector<double> &v, const qdx_grid &g) {
    const long n = grid_points(g);
    std::vector<cplx> a(n);
    for (long i = 0; i < n; ++i) a[i] = cplx(v[2 * i], v[2 * i + 1]);
    fft_grid(a, g, +1);
    for (long i = 0; i < n; ++i) {
        v[2 * i] = a[i].real();
        v[2 * i + 1] = a[i].imag();
    }
}

// OGridPotential::Init with a file (OGridPotential.cpp:36-80), either format
std::vector<double> load_gridop(const std::string &name, bool setzero, double offset, double scale, qdx_grid &g) {
    std::vector<double> v;
    const std::string path = data_name(name, kOpSuffix, -1);
    if (is_stream(path)) {
        const std::vector<StreamEntry> es = read_stream(path);
        v = grid_payload(es.front().payload, 1, g);
    } else {
        const qdx_meta m = file_meta(name);
        g = grid_from_meta(m, name);
        v.resize(grid_points(g));
        read_raw(path, v.data(), v.size() * sizeof(double));
    }
    if (setzero) {
        const double mn = *std::min_element(v.begin(), v.end());
        for (double &x : v) x -= mn;
    }
    if (offset != 0)
        for (double &x : v) x += offset;
    if (scale != 1)
        for (double &x : v) x *= scale;
    return v;
}

// Laser::Init (Laser.cpp:35-50), or a stream (Laser::DeSerialize,
// Laser.cpp:219-235: 64-bit size, dt, values)
std::vector<double> load_laser(const std::string &name, double &dt) {
    const std::string path = data_name(name, kOpSuffix, -1);
    if (is_stream(path)) {
        const std::vector<StreamEntry> es = read_stream(path);
        const std::string &pl = es.front().payload;
        uint64_t n = 0;
        if (pl.size() < 16) fail(QDX_ERR_PARAM, "Laser field has zero size");
        std::memcpy(&n, pl.data(), 8);
        std::memcpy(&dt, pl.data() + 8, 8);
        if (n == 0) fail(QDX_ERR_PARAM, "Laser field has zero size");
        if (!(dt > 0)) fail(QDX_ERR_PARAM, "Laser field got a non-sense dt");
        if (pl.size() < 16 + n * 8) fail(QDX_ERR_IO, "stream payload truncated");
        std::vector<double> v(n);
        std::memcpy(v.data(), pl.data() + 16, n * 8);
        return v;
    }
    const qdx_meta m = file_meta(name);
    if (!m.kv.count("dt")) fail(QDX_ERR_PARAM, "The laser is missing dt");
    if (!m.kv.count("Nt")) fail(QDX_ERR_PARAM, "The laser is missing the number of time steps");
    dt = meta_double(m, "dt", name);
    const long nt = meta_int(m, "Nt", name);
    if (nt < 1) fail(QDX_ERR_PARAM, "Laser field has zero size");
    std::vector<double> v(nt);
    read_raw(data_name(name, kOpSuffix, -1), v.data(), v.size() * sizeof(double));
    return v;
}

Grid grid_of(const qdx_grid &q) {
    Grid g;
    g.nd = q.ndims;
    for (int d = 0; d < q.ndims; ++d) {
        g.N[d] = q.N[d];
        g.xmin[d] = q.xmin[d];
        g.xmax[d] = q.xmax[d];
    }
    return g;
}

std::vector<double> read_state(const std::string &name, int counter, const Grid &g) {
    const qdx_meta m = file_meta(name);
    const std::string cls = m.kv.count("CLASS") ? m.kv.at("CLASS") : "";
    if (cls != "WFGridCartesian") fail(QDX_ERR_INCOMPAT, name + ": class '" + cls + "' is not WFGridCartesian");
    const qdx_grid fg = grid_from_meta(m, name);
    check_same_grid(fg, g, name);
    std::vector<double> v(2 * g.size());
    if (read_raw(data_name(name, kWfSuffix, counter), v.data(), v.size() * sizeof(double)))
        wf_restore(v, fg);   // FileWF::ReadWaveFunction: Restore of compressed data
    return v;
}

// the states of a stream wave function file (FileWF.cpp:112-128): one
// WFGridCartesian, or a WFMultistate entry followed by its states
std::vector<std::vector<double>> read_stream_wf(const std::string &path, qdx_grid &g) {
    const std::vector<StreamEntry> es = read_stream(path);
    size_t first = 0;
    if (es.front().h.cls == "WFMultistate") {
        first = 1;
        if (es.size() < 2) fail(QDX_ERR_PARAM, "Init from stream: Invalid number of states");
    }
    std::vector<std::vector<double>> states;
    for (size_t i = first; i < es.size(); ++i) {
        if (es[i].h.cls != "WFGridCartesian")
            fail(QDX_ERR_INCOMPAT, path + ": wave function class '" + es[i].h.cls + "' is not on the device path");
        qdx_grid gi;
        std::vector<double> v = grid_payload(es[i].payload, 2, gi);
        if (es[i].h.comp != 0) wf_restore(v, gi);
        if (i == first) g = gi;
        else check_same_grid(gi, grid_of(g), path);
        states.push_back(std::move(v));
    }
    return states;
}


}  // namespace

// ---------------------------------------------------------------------------
// meta files

int qdx_meta_create(qdx_meta **out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "NULL argument");
        *out = new qdx_meta();
    });
}

int qdx_meta_read(const char *path, qdx_meta **out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string p = need_str(path, "path");
        auto m = std::make_unique<qdx_meta>();
        if (!read_meta_file(p, *m)) fail(QDX_ERR_IO, "Can not read meta file " + p);
        *out = m.release();
    });
}

int qdx_meta_write(const qdx_meta *m, const char *path) {
    return guarded([&] {
        if (!m) fail(QDX_ERR_PARAM, "NULL argument");
        write_meta_file(*m, need_str(path, "path"));
    });
}

int qdx_meta_size(const qdx_meta *m, int32_t *n) {
    return guarded([&] {
        if (!m || !n) fail(QDX_ERR_PARAM, "NULL argument");
        *n = (int32_t)m->kv.size();
    });
}

static void copy_out(const std::string &s, char *dst, int32_t cap) {
    if (!dst || cap < 1) fail(QDX_ERR_PARAM, "no output buffer");
    if ((int64_t)s.size() + 1 > cap) fail(QDX_ERR_PARAM, "output buffer too small");
    std::memcpy(dst, s.c_str(), s.size() + 1);
}

int qdx_meta_key(const qdx_meta *m, int32_t i, char *key, int32_t cap) {
    return guarded([&] {
        if (!m) fail(QDX_ERR_PARAM, "NULL argument");
        if (i < 0 || i >= (int32_t)m->kv.size()) fail(QDX_ERR_PARAM, "key index out of range");
        auto it = m->kv.begin();
        std::advance(it, i);
        copy_out(it->first, key, cap);
    });
}

int qdx_meta_get(const qdx_meta *m, const char *key, char *value, int32_t cap) {
    return guarded([&] {
        if (!m) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string k = need_str(key, "key");
        auto it = m->kv.find(k);
        if (it == m->kv.end()) fail(QDX_ERR_PARAM, "key '" + k + "' not present");
        copy_out(it->second, value, cap);
    });
}

static void check_key(const std::string &k) {
    if (trim(k).empty() || trim(k) != k || k[0] == '#' || k.find_first_of("=[\n") != std::string::npos)
        fail(QDX_ERR_PARAM, "invalid key '" + k + "'");
}

int qdx_meta_set(qdx_meta *m, const char *key, const char *value) {
    return guarded([&] {
        if (!m) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string k = need_str(key, "key"), v = need_str(value, "value");
        check_key(k);
        if (v.find('\n') != std::string::npos) fail(QDX_ERR_PARAM, "value holds a line break");
        m->kv[k] = trim(v);
    });
}

int qdx_meta_set_int(qdx_meta *m, const char *key, int64_t value) {
    return guarded([&] {
        if (!m) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string k = need_str(key, "key");
        check_key(k);
        m->kv[k] = std::to_string(value);
    });
}

int qdx_meta_set_double(qdx_meta *m, const char *key, double value) {
    return guarded([&] {
        if (!m) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string k = need_str(key, "key");
        check_key(k);
        m->kv[k] = fmt_double(value);
    });
}

int qdx_meta_destroy(qdx_meta *m) {
    return guarded([&] { delete m; });
}

// ---------------------------------------------------------------------------
// raw files

int qdx_file_meta(const char *name, qdx_meta **out) {
    return guarded([&] {
        if (!out) fail(QDX_ERR_PARAM, "NULL argument");
        auto m = std::make_unique<qdx_meta>(file_meta(need_str(name, "name")));
        *out = m.release();
    });
}

int qdx_file_read(const char *name, const char *suffix, int32_t counter, void *dst, int64_t bytes) {
    return guarded([&] {
        if (!dst || bytes <= 0) fail(QDX_ERR_PARAM, "Wrong size");
        read_raw(data_name(need_str(name, "name"), need_str(suffix, "suffix"), counter), dst, (size_t)bytes);
    });
}

int qdx_file_write(const char *name, const char *suffix, int32_t counter, const qdx_meta *meta, const void *src,
                   int64_t bytes) {
    return guarded([&] {
        if (!src || bytes <= 0) fail(QDX_ERR_PARAM, "Wrong size");
        const std::string n = need_str(name, "name");
        if (meta) write_meta_file(*meta, n + kMetaSuffix);
        write_raw(data_name(n, need_str(suffix, "suffix"), counter), src, (size_t)bytes);
    });
}

// ---------------------------------------------------------------------------
// operators, lasers, Propagation.meta

int qdx_gridop_load(const char *name, int32_t setzero, double offset, double scale, qdx_grid *grid, double *data,
                    int64_t cap, int64_t *n) {
    return guarded([&] {
        if (!grid || !n) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string nm = need_str(name, "name");
        if (!data) {   // grid only
            if (is_stream(data_name(nm, kOpSuffix, -1))) {
                load_gridop(nm, false, 0.0, 1.0, *grid);
            } else {
                *grid = grid_from_meta(file_meta(nm), nm);
            }
            *n = grid_points(*grid);
            return;
        }
        qdx_grid g;
        std::vector<double> v = load_gridop(nm, setzero != 0, offset, scale, g);
        if ((int64_t)v.size() > cap) fail(QDX_ERR_PARAM, "output buffer too small");
        std::copy(v.begin(), v.end(), data);
        *grid = g;
        *n = (int64_t)v.size();
    });
}

int qdx_gridop_save(const char *name, const qdx_grid *grid, const double *data) {
    return guarded([&] {
        if (!grid || !data) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string nm = need_str(name, "name");
        if (grid->ndims < 1 || grid->ndims > 4) fail(QDX_ERR_PARAM, "dims out of range");
        qdx_meta m;
        grid_to_meta(*grid, m);
        if (g_io.wf_format == QDX_FORMAT_STREAM) {   // OGridSystem::Serialize
            StreamHeader h;
            h.cls = "OGridPotential";
            h.name = data_name(nm, kOpSuffix, -1);
            h.meta = meta_text(m);
            std::string out = stream_intro();
            put_entry(out, h, grid_payload_out(*grid, data, 1));
            write_raw(h.name, out.data(), out.size());
            return;
        }
        m.kv["CLASS"] = "GridPotential";
        write_meta_file(m, nm + kMetaSuffix);
        write_raw(data_name(nm, kOpSuffix, -1), data, grid_points(*grid) * sizeof(double));
    });
}

int qdx_laser_load(const char *name, double *dt, int64_t *nt, double *data, int64_t cap) {
    return guarded([&] {
        if (!dt || !nt) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string nm = need_str(name, "name");
        if (!data) {
            if (is_stream(data_name(nm, kOpSuffix, -1))) {
                double d = 0;
                *nt = (int64_t)load_laser(nm, d).size();
                *dt = d;
                return;
            }
            const qdx_meta m = file_meta(nm);
            *dt = meta_double(m, "dt", nm);
            *nt = meta_int(m, "Nt", nm);
            return;
        }
        double d = 0;
        std::vector<double> v = load_laser(nm, d);
        if ((int64_t)v.size() > cap) fail(QDX_ERR_PARAM, "output buffer too small");
        std::copy(v.begin(), v.end(), data);
        *dt = d;
        *nt = (int64_t)v.size();
    });
}

int qdx_laser_save(const char *name, double dt, int64_t nt, const double *data) {
    return guarded([&] {
        if (!data || nt < 1) fail(QDX_ERR_PARAM, "Laser field has zero size");
        if (!(dt > 0)) fail(QDX_ERR_PARAM, "Laser field got a non-sense dt");
        const std::string nm = need_str(name, "name");
        qdx_meta m;   // Laser::Dt / Laser::Nt (Laser.cpp:189-204)
        m.kv["dt"] = fmt_double(dt);
        m.kv["Nt"] = std::to_string(nt);
        if (g_io.wf_format == QDX_FORMAT_STREAM) {   // Laser::Serialize (Laser.cpp:206-217)
            StreamHeader h;
            h.cls = "Laser";
            h.name = data_name(nm, kOpSuffix, -1);
            h.meta = meta_text(m);
            std::string pl;
            const uint64_t n64 = (uint64_t)nt;
            pl.append(reinterpret_cast<const char *>(&n64), 8);
            pl.append(reinterpret_cast<const char *>(&dt), 8);
            pl.append(reinterpret_cast<const char *>(data), (size_t)nt * 8);
            std::string out = stream_intro();
            put_entry(out, h, pl);
            write_raw(h.name, out.data(), out.size());
            return;
        }
        m.kv["CLASS"] = "Laser";
        write_meta_file(m, nm + kMetaSuffix);
        write_raw(data_name(nm, kOpSuffix, -1), data, (size_t)nt * sizeof(double));
    });
}

int qdx_propagation_meta(const char *dir, int64_t nt, double dt, const char *wf_base, int32_t wcycle) {
    return guarded([&] {
        const std::string d = dir ? dir : "";
        qdx_meta m;
        m.kv["CLASS"] = "Propagation";
        m.kv["Nt"] = std::to_string(nt);
        m.kv["dt"] = fmt_double(dt);
        m.kv["WFBaseName"] = need_str(wf_base, "wf_base");
        m.kv["Wcycle"] = std::to_string(wcycle);
        write_meta_file(m, d + "Propagation" + kMetaSuffix);
    });
}

// ---------------------------------------------------------------------------
// wave functions and Hamiltonian terms from files

int qdx_wf_load(qdx_ctx *ctx, const char *name, int32_t counter, qdx_wf **out) {
    return guarded([&] {
        if (!ctx || !out) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string base = need_str(name, "name");
        const std::string spath = data_name(base, kWfSuffix, counter);
        if (is_stream(spath)) {
            qdx_grid g{};
            const std::vector<std::vector<double>> st = read_stream_wf(spath, g);
            std::vector<double> all;
            for (const auto &v : st) all.insert(all.end(), v.begin(), v.end());
            qdx_wf *w = nullptr;
            int rc = qdx_wf_create(ctx, &g, (int32_t)st.size(), &w);
            if (rc != QDX_OK) fail(rc, g_err);
            rc = qdx_wf_upload(w, -1, all.data());
            if (rc != QDX_OK) {
                const std::string e = g_err;
                qdx_wf_destroy(w);
                fail(rc, e);
            }
            *out = w;
            return;
        }
        const qdx_meta m = file_meta(base);
        const std::string cls = m.kv.count("CLASS") ? m.kv.at("CLASS") : "";
        if (cls.size() <= 1) fail(QDX_ERR_PARAM, "No CLASS information found. LoadWaveFunctionByMeta failed.");
        std::vector<std::string> names;
        qdx_grid g{};
        if (cls == "WFMultistate") {
            const long states = meta_int(m, "states", base);
            if (states < 1) fail(QDX_ERR_PARAM, "Init from meta file: Invalid number of states");
            for (long i = 0; i < states; ++i) names.push_back(base + "-" + std::to_string(i));
            const qdx_meta m0 = file_meta(names[0]);
            g = grid_from_meta(m0, names[0]);
        } else if (cls == "WFGridCartesian") {
            names.push_back(base);
            g = grid_from_meta(m, base);
        } else {
            fail(QDX_ERR_INCOMPAT, base + ": wave function class '" + cls + "' is not on the device path");
        }
        const Grid grid = grid_of(g);
        std::vector<double> all;
        all.reserve(2 * grid.size() * names.size());
        for (const auto &n : names) {
            std::vector<double> v = read_state(n, counter, grid);
            all.insert(all.end(), v.begin(), v.end());
        }
        qdx_wf *w = nullptr;
        int rc = qdx_wf_create(ctx, &g, (int32_t)names.size(), &w);
        if (rc != QDX_OK) fail(rc, g_err);
        rc = qdx_wf_upload(w, -1, all.data());
        if (rc != QDX_OK) {
            const std::string e = g_err;
            qdx_wf_destroy(w);
            fail(rc, e);
        }
        *out = w;
    });
}

int qdx_wf_save(qdx_wf *wf, const char *name, int32_t counter) {
    return guarded([&] {
        if (!wf) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string base = need_str(name, "name");
        std::vector<double> all(2 * wf->n() * wf->S);
        int rc = qdx_wf_download(wf, -1, all.data());
        if (rc != QDX_OK) fail(rc, g_err);
        const qdx_grid g = to_qdx_grid(wf->grid);
        const bool cmp = g_io.compress != 0;
        const size_t nbytes = (size_t)2 * wf->n() * sizeof(double);
        if (g_io.wf_format == QDX_FORMAT_STREAM) {
            // FileWF::WriteWaveFunction, stream (FileWF.cpp:262-276): a
            // WFMultistate entry (its state count) before the states, every
            // entry but the last marked more_files_follow
            const std::string path = data_name(base, kWfSuffix, counter);
            std::string out = stream_intro();
            StreamHeader h;
            h.name = path;
            h.comp = cmp ? 1 : 0;
            if (wf->S > 1) {
                h.cls = "WFMultistate";
                h.more = true;
                h.meta = "states = " + std::to_string(wf->S) + "\n";
                const uint32_t ns = (uint32_t)wf->S;
                put_entry(out, h, std::string(reinterpret_cast<const char *>(&ns), 4));
            }
            qdx_meta m;
            grid_to_meta(g, m);
            h.cls = "WFGridCartesian";
            h.meta = meta_text(m);
            for (int i = 0; i < wf->S; ++i) {
                const double *src = all.data() + 2 * wf->n() * i;
                std::vector<double> red;
                if (cmp) {
                    red = wf_reduce(src, g, g_io.comptol);
                    src = red.data();
                }
                h.more = i + 1 < wf->S;
                put_entry(out, h, grid_payload_out(g, src, 2));
            }
            write_raw(path, out.data(), out.size());
            return;
        }
        auto write_state = [&](const std::string &n, const double *src) {
            qdx_meta m;   // WFGridCartesian: grid keys + CLASS (FileSingle.h:1303-1314)
            m.kv["CLASS"] = "WFGridCartesian";
            grid_to_meta(g, m);
            write_meta_file(m, n + kMetaSuffix);
            if (cmp) {   // the reduced data, deflated, then the tag (FileSingle.h:1336-1350)
                const std::vector<double> red = wf_reduce(src, g, g_io.comptol);
                std::string z = zlib_deflate(std::string(reinterpret_cast<const char *>(red.data()), nbytes));
                z.append("ZLIBCOM", kTagLen);
                write_raw(data_name(n, kWfSuffix, counter), z.data(), z.size());
                return;
            }
            write_raw(data_name(n, kWfSuffix, counter), src, nbytes);
        };
        if (wf->S == 1) {
            write_state(base, all.data());
            return;
        }
        for (int i = 0; i < wf->S; ++i) write_state(base + "-" + std::to_string(i), all.data() + 2 * wf->n() * i);
        qdx_meta m;   // WFMultistate container (FileWF.cpp:249-255, WFMultistate.cpp:150)
        m.kv["CLASS"] = "WFMultistate";
        m.kv["states"] = std::to_string(wf->S);
        write_meta_file(m, base + kMetaSuffix);
    });
}

int qdx_ham_add_potential_file(qdx_ham *h, int32_t row, int32_t col, const char *name, int32_t setzero,
                               double offset, double scale) {
    return guarded([&] {
        if (!h) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string nm = need_str(name, "name");
        qdx_grid g;
        std::vector<double> v = load_gridop(nm, setzero != 0, offset, scale, g);
        check_same_grid(g, h->grid, nm);
        const int rc = qdx_ham_add_potential(h, row, col, v.data());
        if (rc != QDX_OK) fail(rc, g_err);
    });
}

int qdx_ham_add_dipole_file(qdx_ham *h, int32_t row, int32_t col, const char *mu_name, const char *laser_name,
                            double scale) {
    return guarded([&] {
        if (!h) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string mn = need_str(mu_name, "mu_name"), ln = need_str(laser_name, "laser_name");
        qdx_grid g;
        // OGridDipole::Init reads the array with OGridPotential::Init, which
        // applies `scale` to it (OGridPotential.cpp:71-75), and OLaser::Clock
        // gives the laser the same scale (OLaser.cpp:124-130): QDng's field
        // term is -scale^2 mu(x) E(t)
        std::vector<double> mu = load_gridop(mn, false, 0.0, scale, g);
        check_same_grid(g, h->grid, mn);
        double dt = 0;
        std::vector<double> laser = load_laser(ln, dt);
        if ((int64_t)laser.size() > INT32_MAX) fail(QDX_ERR_PARAM, "laser too long");
        const int rc = qdx_ham_add_dipole(h, row, col, mu.data(), laser.data(), (int32_t)laser.size(), scale);
        if (rc != QDX_OK) fail(rc, g_err);
    });
}

// ---------------------------------------------------------------------------
// output format

int qdx_io_set_format(int32_t wf_format, int32_t compress, double comptol) {
    return guarded([&] {
        if (wf_format != QDX_FORMAT_BINARY && wf_format != QDX_FORMAT_STREAM)
            fail(QDX_ERR_PARAM, "Unrecognized file wave function file format");
        g_io.wf_format = wf_format;
        g_io.compress = compress != 0;
        g_io.comptol = comptol > 0 ? comptol : 1e-6;
    });
}

int qdx_io_get_format(int32_t *wf_format, int32_t *compress, double *comptol) {
    return guarded([&] {
        if (!wf_format || !compress || !comptol) fail(QDX_ERR_PARAM, "NULL argument");
        *wf_format = g_io.wf_format;
        *compress = g_io.compress;
        *comptol = g_io.comptol;
    });
}

int qdx_file_format(const char *path, int32_t *format) {
    return guarded([&] {
        if (!format) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string p = need_str(path, "path");
        std::ifstream f(p, std::ios::binary);
        if (!f.is_open()) fail(QDX_ERR_IO, "Can not open file for reading: " + p);
        *format = is_stream(p) ? QDX_FORMAT_STREAM : QDX_FORMAT_BINARY;
    });
}
