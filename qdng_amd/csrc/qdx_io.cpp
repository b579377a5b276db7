// QDng's binary file format on the host: key/value meta files
// (ParamContainer), raw data files with sequence naming (FileSingle, binary
// format), operator and laser files, Propagation.meta, and wave functions /
// Hamiltonian terms loaded from them.  The C interface is include/qdx_io.h.
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

// FileSingle::DetectCompressionInPayload + ReadDataFromRaw (uncompressed)
void read_raw(const std::string &path, void *dst, size_t bytes) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) fail(QDX_ERR_IO, "Can not open binary file for reading: " + path);
    f.seekg(0, std::ios::end);
    const std::streamoff len = f.tellg();
    if (len < kTagLen) fail(QDX_ERR_IO, "Binary file truncated: " + path);
    char tag[kTagLen];
    f.seekg(len - kTagLen);
    f.read(tag, kTagLen);
    if (std::strncmp(tag, "ZLIBCOM", kTagLen) == 0 || std::strncmp(tag, "BZIPCOM", kTagLen) == 0)
        fail(QDX_ERR_INCOMPAT, "compressed binary file (" + std::string(tag) + "): not supported: " + path);
    f.seekg(0);
    f.read(static_cast<char *>(dst), (std::streamsize)bytes);
    if ((size_t)f.gcount() != bytes) fail(QDX_ERR_IO, "Can not read binary file: " + path);
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

// OGridPotential::Init with a file (OGridPotential.cpp:36-80)
std::vector<double> load_gridop(const std::string &name, bool setzero, double offset, double scale, qdx_grid &g) {
    const qdx_meta m = file_meta(name);
    g = grid_from_meta(m, name);
    std::vector<double> v(grid_points(g));
    read_raw(data_name(name, kOpSuffix, -1), v.data(), v.size() * sizeof(double));
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

// Laser::Init (Laser.cpp:35-50)
std::vector<double> load_laser(const std::string &name, double &dt) {
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

std::vector<double> read_state(const std::string &name, int counter, const Grid &g) {
    const qdx_meta m = file_meta(name);
    const std::string cls = m.kv.count("CLASS") ? m.kv.at("CLASS") : "";
    if (cls != "WFGridCartesian") fail(QDX_ERR_INCOMPAT, name + ": class '" + cls + "' is not WFGridCartesian");
    check_same_grid(grid_from_meta(m, name), g, name);
    std::vector<double> v(2 * g.size());
    read_raw(data_name(name, kWfSuffix, counter), v.data(), v.size() * sizeof(double));
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
            *grid = grid_from_meta(file_meta(nm), nm);
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
        m.kv["CLASS"] = "GridPotential";
        grid_to_meta(*grid, m);
        write_meta_file(m, nm + kMetaSuffix);
        write_raw(data_name(nm, kOpSuffix, -1), data, grid_points(*grid) * sizeof(double));
    });
}

int qdx_laser_load(const char *name, double *dt, int64_t *nt, double *data, int64_t cap) {
    return guarded([&] {
        if (!dt || !nt) fail(QDX_ERR_PARAM, "NULL argument");
        const std::string nm = need_str(name, "name");
        if (!data) {
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
        m.kv["CLASS"] = "Laser";
        m.kv["dt"] = fmt_double(dt);
        m.kv["Nt"] = std::to_string(nt);
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
        auto write_state = [&](const std::string &n, const double *src) {
            qdx_meta m;   // WFGridCartesian: grid keys + CLASS (FileSingle.h:1303-1314)
            m.kv["CLASS"] = "WFGridCartesian";
            grid_to_meta(g, m);
            write_meta_file(m, n + kMetaSuffix);
            write_raw(data_name(n, kWfSuffix, counter), src, (size_t)2 * wf->n() * sizeof(double));
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
        std::vector<double> mu = load_gridop(mn, false, 0.0, 1.0, g);
        check_same_grid(g, h->grid, mn);
        double dt = 0;
        std::vector<double> laser = load_laser(ln, dt);
        if ((int64_t)laser.size() > INT32_MAX) fail(QDX_ERR_PARAM, "laser too long");
        const int rc = qdx_ham_add_dipole(h, row, col, mu.data(), laser.data(), (int32_t)laser.size(), scale);
        if (rc != QDX_OK) fail(rc, g_err);
    });
}
