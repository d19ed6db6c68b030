// Mesh2HRTF / NumCalc project input (NC.inp and its node/element files).
//
// Restates the reference's parser and loader:
//   parse_nc_input / parse_nc_input_string  crates/sonium-bem/src/core/io/nc_format.rs:205-663
//   line helpers                            nc_format.rs:665-692
//   load_nc_nodes / load_nc_elements        nc_format.rs:694-781
//   load_project, assemble_mesh,
//   apply_boundary_specs, expand_frequency_curve
//                                           crates/sonium-bem/src/numcalc/project.rs:78-238
// Host code only: it produces the arrays that sbh_mesh_create_oriented and
// the field evaluation take.  Tokens are parsed with Rust's rules for i32,
// usize and f64 (optional sign, decimal digits; f64 also accepts exponents,
// "inf", "infinity" and "nan" in any case; no hexadecimal), so accepted and
// rejected records match the reference's.

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

using sbh::Error;

namespace {

struct BoundarySpec {
    uint64_t start = 0, end = 0;
    std::string type;
    double re = 0.0, im = 0.0;
    int32_t curve_re = -1, curve_im = -1;
};

struct Source {
    double v[3] = {0.0, 0.0, 0.0};
    double re = 0.0, im = 0.0;
    int32_t curve_re = -1, curve_im = -1;
};

struct Config {
    std::string version, description;
    std::vector<int32_t> control_i;
    std::vector<double> control_ii;
    std::vector<double> frequency_curve;  // (time, frequency, phase) triples
    int32_t main_i[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double main_ii[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<int32_t> main_iii;
    double speed = 343.0, density = 1.21, reference_pressure = 1.0;
    std::vector<double> main_iv_reserved = {0.0, 0.0, 0.0, 0.0};
    std::vector<std::string> node_files, element_files;
    bool has_symmetry = false;
    int32_t symmetry_flags[3] = {0, 0, 0};
    double symmetry_origin[3] = {0.0, 0.0, 0.0};
    std::vector<BoundarySpec> bcs;
    std::vector<Source> plane_waves, point_sources;
    std::string base_dir;
};

Error parse_error(size_t line, const std::string& msg) {
    return Error(SBH_ERR_INVALID_ARGUMENT, "parse error at line " + std::to_string(line) + ": " + msg);
}

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

bool starts_with(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }
bool contains(const std::string& s, const char* p) { return s.find(p) != std::string::npos; }

std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream in(s);
    std::string t;
    while (in >> t) out.push_back(t);
    return out;
}

// Rust's integer grammar: [+-]?[0-9]+ (unsigned types: [+]?[0-9]+), in range.
bool parse_i64_rust(const std::string& t, bool allow_minus, long long lo, long long hi, long long* out) {
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) {
        neg = t[i] == '-';
        if (neg && !allow_minus) return false;
        ++i;
    }
    if (i == t.size()) return false;
    unsigned long long v = 0;
    for (; i < t.size(); ++i) {
        if (t[i] < '0' || t[i] > '9') return false;
        if (v > (std::numeric_limits<unsigned long long>::max() - 9) / 10) return false;
        v = v * 10 + (unsigned)(t[i] - '0');
    }
    if (neg) {
        if (v > (unsigned long long)std::numeric_limits<long long>::max() + 1ull) return false;
        long long s = v == (unsigned long long)std::numeric_limits<long long>::max() + 1ull
                          ? std::numeric_limits<long long>::min()
                          : -(long long)v;
        if (s < lo) return false;
        *out = s;
        return true;
    }
    if (v > (unsigned long long)hi) return false;
    *out = (long long)v;
    return true;
}

bool parse_i32(const std::string& t, int32_t* out) {
    long long v;
    if (!parse_i64_rust(t, true, std::numeric_limits<int32_t>::min(), std::numeric_limits<int32_t>::max(), &v))
        return false;
    *out = (int32_t)v;
    return true;
}

bool parse_usize(const std::string& t, uint64_t* out) {
    long long v;
    if (!parse_i64_rust(t, false, 0, std::numeric_limits<long long>::max(), &v)) return false;
    *out = (uint64_t)v;
    return true;
}

// Rust's f64 grammar: [+-]?(inf|infinity|nan) (any case) or
// [+-]?(digits[.digits?]|.digits)([eE][+-]?digits)?
bool parse_f64(const std::string& t, double* out) {
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) {
        neg = t[i] == '-';
        ++i;
    }
    std::string rest = t.substr(i);
    std::string low = rest;
    for (char& c : low) c = (char)std::tolower((unsigned char)c);
    if (low == "inf" || low == "infinity") {
        *out = neg ? -INFINITY : INFINITY;
        return true;
    }
    if (low == "nan") {
        *out = std::numeric_limits<double>::quiet_NaN();
        return true;
    }
    size_t j = 0, digits = 0;
    while (j < rest.size() && std::isdigit((unsigned char)rest[j])) ++j, ++digits;
    if (j < rest.size() && rest[j] == '.') {
        ++j;
        while (j < rest.size() && std::isdigit((unsigned char)rest[j])) ++j, ++digits;
    }
    if (digits == 0) return false;
    if (j < rest.size() && (rest[j] == 'e' || rest[j] == 'E')) {
        ++j;
        if (j < rest.size() && (rest[j] == '+' || rest[j] == '-')) ++j;
        size_t ed = 0;
        while (j < rest.size() && std::isdigit((unsigned char)rest[j])) ++j, ++ed;
        if (ed == 0) return false;
    }
    if (j != rest.size()) return false;
    errno = 0;
    double v = std::strtod(t.c_str(), nullptr);  // decimal only, validated above
    *out = v;
    return true;
}

// nc_format.rs:638-648 parse_int_line_checked
std::vector<int32_t> int_line_checked(const std::string& line, size_t lineno) {
    std::vector<int32_t> out;
    for (const std::string& t : split_ws(line)) {
        int32_t v;
        if (!parse_i32(t, &v)) throw parse_error(lineno, "invalid integer token \"" + t + "\"");
        out.push_back(v);
    }
    return out;
}

// nc_format.rs:662-671 parse_float_line_checked
std::vector<double> float_line_checked(const std::string& line, size_t lineno) {
    std::vector<double> out;
    for (const std::string& t : split_ws(line)) {
        double v;
        if (!parse_f64(t, &v)) throw parse_error(lineno, "invalid numeric token \"" + t + "\"");
        out.push_back(v);
    }
    return out;
}

// nc_format.rs:650-660 parse_float_line: tokens that do not parse are dropped.
std::vector<double> float_line(const std::string& line) {
    std::vector<double> out;
    for (const std::string& t : split_ws(line)) {
        double v;
        if (parse_f64(t, &v)) out.push_back(v);
    }
    return out;
}

// Rust `as i32` on f64: truncation toward zero, saturating, NaN -> 0.
int32_t f64_as_i32(double v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return std::numeric_limits<int32_t>::max();
    if (v <= -2147483648.0) return std::numeric_limits<int32_t>::min();
    return (int32_t)v;
}

std::string join_path(const std::string& base, const std::string& rel) {
    if (!rel.empty() && rel[0] == '/') return rel;
    if (base.empty() || base == ".") return base.empty() ? rel : "./" + rel;
    return base.back() == '/' ? base + rel : base + "/" + rel;
}

// nc_format.rs:673-690 parse_boundary_line
bool boundary_line(const std::string& line, BoundarySpec* bc) {
    std::vector<std::string> p = split_ws(line);
    if (p.size() < 9 || p[0] != "ELEM" || p[2] != "TO") return false;
    return parse_usize(p[1], &bc->start) && parse_usize(p[3], &bc->end) && (bc->type = p[4], true) &&
           parse_f64(p[5], &bc->re) && parse_i32(p[6], &bc->curve_re) && parse_f64(p[7], &bc->im) &&
           parse_i32(p[8], &bc->curve_im);
}

// nc_format.rs:692-722 parse_plane_wave_line / parse_point_source_line
bool source_line(const std::string& line, Source* s) {
    std::vector<double> v = float_line(line);
    if (v.size() < 8) return false;
    s->v[0] = v[1];
    s->v[1] = v[2];
    s->v[2] = v[3];
    s->re = v[4];
    s->curve_re = f64_as_i32(v[5]);
    s->im = v[6];
    s->curve_im = f64_as_i32(v[7]);
    return true;
}

std::vector<std::string> split_lines(const std::string& content) {
    std::vector<std::string> lines;
    size_t b = 0;
    while (b < content.size()) {
        size_t e = content.find('\n', b);
        if (e == std::string::npos) e = content.size();
        std::string l = content.substr(b, e - b);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        lines.push_back(l);
        b = e + 1;
    }
    return lines;
}

// nc_format.rs:215-636 parse_nc_input_string.  Keyword records are recognised
// by the raw text of the line before them, as the reference does.
Config parse_config(const std::string& content, const std::string& base_dir) {
    Config c;
    c.base_dir = base_dir;
    const std::vector<std::string> lines = split_lines(content);
    auto prev = [&](size_t i) -> const std::string& {
        static const std::string empty;
        return i > 0 ? lines[i - 1] : empty;
    };
    size_t i = 0;
    bool saw_nodes = false, saw_elements = false, saw_end = false;
    while (i < lines.size()) {
        const std::string line = trim(lines[i]);
        if (line.empty() || starts_with(line, "##") || starts_with(line, "#")) {
            ++i;
            continue;
        }
        if (starts_with(line, "Mesh2HRTF")) {
            c.version = line;
            ++i;
            continue;
        }
        if (c.version.empty()) {
            ++i;
            continue;
        }
        const std::string& pl = prev(i);
        if (i > 0 && contains(pl, "Controlparameter I") && !contains(pl, "Controlparameter II")) {
            c.control_i = int_line_checked(line, i + 1);
            ++i;
            continue;
        }
        if (i > 0 && contains(pl, "Controlparameter II")) {
            c.control_ii = float_line_checked(line, i + 1);
            ++i;
            continue;
        }
        if (i > 0 && contains(pl, "Frequency Curve")) {
            std::vector<int32_t> header = int_line_checked(line, i + 1);
            if (header.size() < 2) throw parse_error(i + 1, "frequency curve header requires a point count");
            if (header[1] < 0) throw parse_error(i + 1, "frequency curve point count must be non-negative");
            const size_t npts = (size_t)header[1];
            ++i;
            for (size_t p = 0; p < npts; ++p) {
                if (i >= lines.size())
                    throw parse_error(i + 1, "frequency curve is shorter than its declared point count");
                std::vector<double> v = float_line_checked(lines[i], i + 1);
                if (v.size() < 3)
                    throw parse_error(i + 1, "frequency curve point requires time, frequency, and phase");
                c.frequency_curve.insert(c.frequency_curve.end(), {v[0], v[1], v[2]});
                ++i;
            }
            continue;
        }
        if (i > 0 && contains(pl, "Main Parameters I") && !contains(pl, "Main Parameters II") &&
            !contains(pl, "Main Parameters III") && !contains(pl, "Main Parameters IV")) {
            std::vector<int32_t> v = int_line_checked(line, i + 1);
            if (v.size() < 9) throw parse_error(i + 1, "Main Parameters I requires nine integer values");
            for (int q = 0; q < 9; ++q) c.main_i[q] = v[q];
            ++i;
            continue;
        }
        if (i > 0 && contains(pl, "Main Parameters II") && !contains(pl, "Main Parameters III") &&
            !contains(pl, "Main Parameters IV")) {
            std::vector<double> v = float_line_checked(line, i + 1);
            if (v.size() < 7) throw parse_error(i + 1, "Main Parameters II requires seven values");
            for (int q = 0; q < 7; ++q) c.main_ii[q] = v[q];
            ++i;
            continue;
        }
        if (i > 0 && contains(pl, "Main Parameters III") && !contains(pl, "Main Parameters IV")) {
            c.main_iii = int_line_checked(line, i + 1);
            ++i;
            continue;
        }
        if (i > 0 && contains(pl, "Main Parameters IV")) {
            std::vector<double> v = float_line_checked(line, i + 1);
            if (v.size() < 3)
                throw parse_error(i + 1, "Main Parameters IV requires speed, density, and reference pressure");
            c.speed = v[0];
            c.density = v[1];
            c.reference_pressure = v[2];
            c.main_iv_reserved.assign(v.begin() + 3, v.end());
            ++i;
            continue;
        }
        if (line == "NODES" || line == "ELEMENTS") {
            const bool nodes = line == "NODES";
            (nodes ? saw_nodes : saw_elements) = true;
            ++i;
            while (i < lines.size()) {
                const std::string l = trim(lines[i]);
                if (starts_with(l, "##") || l.empty()) break;
                if (!starts_with(l, "#")) (nodes ? c.node_files : c.element_files).push_back(join_path(base_dir, l));
                ++i;
            }
            continue;
        }
        if (line == "SYMMETRY") {
            ++i;
            if (i < lines.size()) {
                const std::string fl = trim(lines[i]);
                if (!starts_with(fl, "#")) {
                    std::vector<int32_t> flags = int_line_checked(fl, i + 1);
                    if (flags.size() < 3) throw parse_error(i + 1, "symmetry flags require three integer values");
                    ++i;
                    if (i < lines.size()) {
                        std::vector<double> o = float_line_checked(trim(lines[i]), i + 1);
                        if (o.size() < 3) throw parse_error(i + 1, "symmetry origin requires three values");
                        c.has_symmetry = true;
                        for (int q = 0; q < 3; ++q) {
                            c.symmetry_flags[q] = flags[q] != 0 ? 1 : 0;
                            c.symmetry_origin[q] = o[q];
                        }
                        ++i;
                    }
                }
            }
            continue;
        }
        if (line == "BOUNDARY") {
            ++i;
            while (i < lines.size()) {
                const std::string l = trim(lines[i]);
                if (starts_with(l, "##") || l == "RETU") {
                    ++i;
                    break;
                }
                if (starts_with(l, "#") || l.empty()) {
                    ++i;
                    continue;
                }
                BoundarySpec bc;
                if (!boundary_line(l, &bc)) throw parse_error(i + 1, "invalid boundary condition record");
                c.bcs.push_back(bc);
                ++i;
            }
            continue;
        }
        if (line == "PLANE WAVES" || line == "POINT SOURCES") {
            const bool plane = line == "PLANE WAVES";
            ++i;
            while (i < lines.size()) {
                const std::string l = trim(lines[i]);
                if (starts_with(l, "##") || l.empty()) break;
                if (!starts_with(l, "#")) {
                    Source s;
                    if (!source_line(l, &s))
                        throw parse_error(i + 1, plane ? "invalid plane-wave source record" : "invalid point-source record");
                    (plane ? c.plane_waves : c.point_sources).push_back(s);
                }
                ++i;
            }
            continue;
        }
        if (line == "END") {
            saw_end = true;
            break;
        }
        ++i;
    }
    auto missing = [](const char* what) {
        return Error(SBH_ERR_INVALID_ARGUMENT, std::string("missing required section: ") + what);
    };
    if (c.version.empty()) throw missing("Mesh2HRTF header");
    if (!saw_nodes) throw missing("NODES");
    if (!saw_elements) throw missing("ELEMENTS");
    if (!saw_end) throw missing("END");
    return c;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(SBH_ERR_INVALID_ARGUMENT, "IO error: cannot read " + path);
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

bool first_line_is_count(const std::vector<std::string>& lines) {
    uint64_t v;
    return !lines.empty() && parse_usize(trim(lines[0]), &v);
}

// nc_format.rs:694-722 load_nc_nodes
void load_nodes(const std::string& path, std::vector<double>* nodes) {
    const std::vector<std::string> lines = split_lines(read_file(path));
    for (size_t i = first_line_is_count(lines) ? 1 : 0; i < lines.size(); ++i) {
        std::vector<double> v = float_line(lines[i]);
        if (v.size() >= 4)
            nodes->insert(nodes->end(), {v[1], v[2], v[3]});
        else if (v.size() >= 3)
            nodes->insert(nodes->end(), {v[0], v[1], v[2]});
    }
}

struct RawElement {
    std::vector<int64_t> conn;
};

// nc_format.rs:724-781 load_nc_elements: integer tokens of each line; with at
// least 4 of them, the connectivity is the non-negative run after the id.
void load_elements(const std::string& path, std::vector<RawElement>* out) {
    const std::vector<std::string> lines = split_lines(read_file(path));
    for (size_t i = first_line_is_count(lines) ? 1 : 0; i < lines.size(); ++i) {
        std::vector<int32_t> v;
        for (const std::string& t : split_ws(lines[i])) {
            int32_t x;
            if (parse_i32(t, &x)) v.push_back(x);
        }
        if (v.size() < 4) continue;
        RawElement e;
        for (size_t q = 1; q < v.size() && v[q] >= 0; ++q) e.conn.push_back(v[q]);
        out->push_back(e);
    }
}

}  // namespace

struct sbh_nc_project {
    Config config;
    bool mesh_loaded = false;
    std::vector<double> nodes;              // V x 3
    std::vector<int64_t> conn;              // E x 4, -1 padded (Tri3: 3 entries)
    std::vector<int32_t> conn_count;        // entries per element
    std::vector<uint8_t> is_evaluation;     // E
    std::vector<int32_t> bc_type;           // E: 0 velocity, 1 pressure
    std::vector<sbh_c64> bc_value;          // E
    std::vector<sbh_c64> admittance;        // E
    std::vector<uint8_t> has_admittance;    // E
    std::vector<int64_t> evaluation_indices;
    std::vector<double> frequencies;
};

namespace {

// project.rs:115-169 assemble_mesh
void assemble_mesh(sbh_nc_project* p) {
    const Config& c = p->config;
    const size_t n_files = std::min(c.node_files.size(), c.element_files.size());
    const size_t n_object = std::min((size_t)std::max<int32_t>(c.main_i[3], 0), n_files);
    std::vector<size_t> counts;
    for (size_t f = 0; f < n_files; ++f) {
        const size_t before = p->nodes.size() / 3;
        load_nodes(c.node_files[f], &p->nodes);
        counts.push_back(p->nodes.size() / 3 - before);
    }
    size_t offset = 0;
    for (size_t f = 0; f < n_files; ++f) {
        std::vector<RawElement> chunk;
        load_elements(c.element_files[f], &chunk);
        const bool eval = f >= n_object;
        for (RawElement& e : chunk) {
            int64_t row[4] = {-1, -1, -1, -1};
            const size_t cnt = e.conn.size();
            for (size_t q = 0; q < cnt && q < 4; ++q) row[q] = e.conn[q] + (int64_t)offset;
            if (eval) p->evaluation_indices.push_back((int64_t)p->is_evaluation.size());
            p->conn.insert(p->conn.end(), row, row + 4);
            p->conn_count.push_back((int32_t)cnt);
            p->is_evaluation.push_back(eval ? 1 : 0);
        }
        offset += counts[f];
    }
    const size_t ne = p->is_evaluation.size();
    p->bc_type.assign(ne, 0);
    p->bc_value.assign(ne, sbh_c64{0.0, 0.0});
    p->admittance.assign(ne, sbh_c64{0.0, 0.0});
    p->has_admittance.assign(ne, 0);
}

// project.rs:171-205 apply_boundary_specs over all elements in file order.
void apply_boundary_specs(sbh_nc_project* p) {
    const uint64_t ne = p->is_evaluation.size();
    for (const BoundarySpec& s : p->config.bcs) {
        if (s.start > s.end)
            throw Error(SBH_ERR_INVALID_ARGUMENT, "boundary spec covers element " + std::to_string(s.start) +
                                                      " but mesh only has " + std::to_string(s.end) + " elements");
        if (s.end >= ne)
            throw Error(SBH_ERR_INVALID_ARGUMENT, "boundary spec covers element " + std::to_string(s.end) +
                                                      " but mesh only has " + std::to_string(ne) + " elements");
        int32_t type;
        bool adm = false;
        if (s.type == "VELO")
            type = 0;
        else if (s.type == "PRES")
            type = 1;
        else if (s.type == "ADMI")
            type = 0, adm = true;
        else
            throw Error(SBH_ERR_UNSUPPORTED, "unsupported boundary condition keyword: " + s.type);
        for (uint64_t e = s.start; e <= s.end; ++e) {
            p->bc_type[e] = type;
            p->has_admittance[e] = adm ? 1 : 0;
            p->admittance[e] = adm ? sbh_c64{s.re, s.im} : sbh_c64{0.0, 0.0};
            p->bc_value[e] = adm ? sbh_c64{0.0, 0.0} : sbh_c64{s.re, s.im};
        }
    }
}

// project.rs:207-225 expand_frequency_curve: positive finite, sorted, unique.
void expand_frequencies(sbh_nc_project* p) {
    std::vector<double> f;
    for (size_t i = 0; 3 * i + 1 < p->config.frequency_curve.size(); ++i) {
        const double v = p->config.frequency_curve[3 * i + 1];
        if (std::isfinite(v) && v > 0.0) f.push_back(v);
    }
    std::sort(f.begin(), f.end());
    f.erase(std::unique(f.begin(), f.end()), f.end());
    p->frequencies = f;
}

std::string parent_dir(const std::string& path) {
    const size_t s = path.find_last_of('/');
    if (s == std::string::npos) return ".";
    return s == 0 ? "/" : path.substr(0, s);
}

void copy_string(const std::string& s, char* buf, int64_t size) {
    SBH_REQUIRE(buf && size > 0, "buffer is null or empty");
    const size_t n = std::min<size_t>(s.size(), (size_t)size - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

}  // namespace

extern "C" {

sbh_status sbh_nc_parse(const char* content, const char* base_dir, sbh_nc_project** out) {
    return sbh::guard([&] {
        SBH_REQUIRE(content && out, "content/out is null");
        std::unique_ptr<sbh_nc_project> p(new sbh_nc_project());
        p->config = parse_config(content, base_dir ? base_dir : ".");
        expand_frequencies(p.get());
        *out = p.release();
    });
}

sbh_status sbh_nc_project_load(const char* nc_inp_path, sbh_nc_project** out) {
    return sbh::guard([&] {
        SBH_REQUIRE(nc_inp_path && out, "path/out is null");
        const std::string path(nc_inp_path);
        std::unique_ptr<sbh_nc_project> p(new sbh_nc_project());
        p->config = parse_config(read_file(path), parent_dir(path));
        assemble_mesh(p.get());
        apply_boundary_specs(p.get());
        expand_frequencies(p.get());
        p->mesh_loaded = true;
        *out = p.release();
    });
}

void sbh_nc_project_destroy(sbh_nc_project* p) { delete p; }

sbh_status sbh_nc_project_info(const sbh_nc_project* p, sbh_nc_info* info) {
    return sbh::guard([&] {
        SBH_REQUIRE(p && info, "project/info is null");
        const Config& c = p->config;
        sbh_nc_info r;
        std::memset(&r, 0, sizeof(r));
        r.num_nodes = (int64_t)p->nodes.size() / 3;
        r.num_elements = (int64_t)p->is_evaluation.size();
        r.num_evaluation = (int64_t)p->evaluation_indices.size();
        r.num_frequencies = (int64_t)p->frequencies.size();
        r.num_curve_points = (int64_t)c.frequency_curve.size() / 3;
        r.num_plane_waves = (int64_t)c.plane_waves.size();
        r.num_point_sources = (int64_t)c.point_sources.size();
        r.num_boundary_specs = (int64_t)c.bcs.size();
        r.num_node_files = (int64_t)c.node_files.size();
        r.num_element_files = (int64_t)c.element_files.size();
        r.mesh_loaded = p->mesh_loaded ? 1 : 0;
        for (int q = 0; q < 9; ++q) r.main_params_i[q] = c.main_i[q];
        for (int q = 0; q < 7; ++q) r.main_params_ii[q] = c.main_ii[q];
        r.speed_of_sound = c.speed;
        r.density = c.density;
        r.reference_pressure = c.reference_pressure;
        r.has_symmetry = c.has_symmetry ? 1 : 0;
        for (int q = 0; q < 3; ++q) {
            r.symmetry_flags[q] = c.symmetry_flags[q];
            r.symmetry_origin[q] = c.symmetry_origin[q];
        }
        *info = r;
    });
}

sbh_status sbh_nc_project_string(const sbh_nc_project* p, int which, int64_t index, char* buf, int64_t size) {
    return sbh::guard([&] {
        SBH_REQUIRE(p, "project is null");
        const Config& c = p->config;
        switch (which) {
            case SBH_NC_VERSION: copy_string(c.version, buf, size); break;
            case SBH_NC_DESCRIPTION: copy_string(c.description, buf, size); break;
            case SBH_NC_NODE_FILE:
                SBH_REQUIRE(index >= 0 && index < (int64_t)c.node_files.size(), "node file index out of range");
                copy_string(c.node_files[index], buf, size);
                break;
            case SBH_NC_ELEMENT_FILE:
                SBH_REQUIRE(index >= 0 && index < (int64_t)c.element_files.size(), "element file index out of range");
                copy_string(c.element_files[index], buf, size);
                break;
            case SBH_NC_BOUNDARY_TYPE:
                SBH_REQUIRE(index >= 0 && index < (int64_t)c.bcs.size(), "boundary spec index out of range");
                copy_string(c.bcs[index].type, buf, size);
                break;
            default: throw Error(SBH_ERR_INVALID_ARGUMENT, "unknown string selector");
        }
    });
}

sbh_status sbh_nc_project_values(const sbh_nc_project* p, int which, double* out, int64_t capacity, int64_t* count) {
    return sbh::guard([&] {
        SBH_REQUIRE(p && count, "project/count is null");
        const Config& c = p->config;
        std::vector<double> v;
        switch (which) {
            case SBH_NC_CONTROL_I: v.assign(c.control_i.begin(), c.control_i.end()); break;
            case SBH_NC_CONTROL_II: v = c.control_ii; break;
            case SBH_NC_MAIN_III: v.assign(c.main_iii.begin(), c.main_iii.end()); break;
            case SBH_NC_MAIN_IV_RESERVED: v = c.main_iv_reserved; break;
            case SBH_NC_FREQUENCY_CURVE: v = c.frequency_curve; break;
            case SBH_NC_FREQUENCIES: v = p->frequencies; break;
            case SBH_NC_BOUNDARY_SPECS:
                for (const BoundarySpec& b : c.bcs)
                    v.insert(v.end(), {(double)b.start, (double)b.end, b.re, (double)b.curve_re, b.im,
                                       (double)b.curve_im});
                break;
            case SBH_NC_PLANE_WAVES:
            case SBH_NC_POINT_SOURCES:
                for (const Source& s : which == SBH_NC_PLANE_WAVES ? c.plane_waves : c.point_sources)
                    v.insert(v.end(), {s.v[0], s.v[1], s.v[2], s.re, (double)s.curve_re, s.im, (double)s.curve_im});
                break;
            default: throw Error(SBH_ERR_INVALID_ARGUMENT, "unknown value selector");
        }
        *count = (int64_t)v.size();
        if (out) {
            SBH_REQUIRE(capacity >= (int64_t)v.size(), "output buffer too small");
            std::copy(v.begin(), v.end(), out);
        }
    });
}

sbh_status sbh_nc_project_mesh(const sbh_nc_project* p, double* nodes, int64_t* connectivity, int32_t* counts,
                               uint8_t* is_evaluation, int32_t* bc_type, sbh_c64* bc_value, sbh_c64* admittance,
                               uint8_t* has_admittance) {
    return sbh::guard([&] {
        SBH_REQUIRE(p, "project is null");
        SBH_REQUIRE(p->mesh_loaded, "the project was parsed without its mesh files");
        if (nodes) std::copy(p->nodes.begin(), p->nodes.end(), nodes);
        if (connectivity) std::copy(p->conn.begin(), p->conn.end(), connectivity);
        if (counts) std::copy(p->conn_count.begin(), p->conn_count.end(), counts);
        if (is_evaluation) std::copy(p->is_evaluation.begin(), p->is_evaluation.end(), is_evaluation);
        if (bc_type) std::copy(p->bc_type.begin(), p->bc_type.end(), bc_type);
        if (bc_value) std::copy(p->bc_value.begin(), p->bc_value.end(), bc_value);
        if (admittance) std::copy(p->admittance.begin(), p->admittance.end(), admittance);
        if (has_admittance) std::copy(p->has_admittance.begin(), p->has_admittance.end(), has_admittance);
    });
}

sbh_status sbh_nc_project_evaluation_centers(const sbh_nc_project* p, double* out) {
    return sbh::guard([&] {
        SBH_REQUIRE(p && out, "project/out is null");
        SBH_REQUIRE(p->mesh_loaded, "the project was parsed without its mesh files");
        const int64_t nn = (int64_t)p->nodes.size() / 3;
        for (size_t k = 0; k < p->evaluation_indices.size(); ++k) {
            const int64_t e = p->evaluation_indices[k];
            const int32_t cnt = p->conn_count[e];
            SBH_REQUIRE(cnt >= 1 && cnt <= 4, "evaluation element has no usable nodes");
            // compute_element_center (element.rs): mean of the element's nodes
            double c[3] = {0.0, 0.0, 0.0};
            for (int32_t q = 0; q < cnt; ++q) {
                const int64_t node = p->conn[4 * e + q];
                SBH_REQUIRE(node >= 0 && node < nn, "evaluation element node index out of range");
                for (int d = 0; d < 3; ++d) c[d] += p->nodes[3 * node + d];
            }
            for (int d = 0; d < 3; ++d) out[3 * k + d] = c[d] / (double)cnt;
        }
    });
}

sbh_status sbh_nc_project_surface_mesh(const sbh_nc_project* p, sbh_context* ctx, sbh_mesh** out) {
    return sbh::guard([&] {
        SBH_REQUIRE(p && ctx && out, "project/ctx/out is null");
        SBH_REQUIRE(p->mesh_loaded, "the project was parsed without its mesh files");
        std::vector<int64_t> tris;
        std::vector<int32_t> bt;
        std::vector<sbh_c64> bv, adm;
        std::vector<uint8_t> has;
        for (size_t e = 0; e < p->is_evaluation.size(); ++e) {
            if (p->is_evaluation[e]) continue;
            if (p->conn_count[e] != 3)
                throw Error(SBH_ERR_UNSUPPORTED, "surface element " + std::to_string(e) +
                                                     " is not a Tri3 panel (only Tri3 is supported)");
            tris.insert(tris.end(), &p->conn[4 * e], &p->conn[4 * e] + 3);
            bt.push_back(p->bc_type[e]);
            bv.push_back(p->bc_value[e]);
            adm.push_back(p->admittance[e]);
            has.push_back(p->has_admittance[e]);
        }
        SBH_REQUIRE(!bt.empty(), "the project has no surface elements");
        const sbh_status st = sbh_mesh_create_oriented(ctx, p->nodes.data(), (int64_t)p->nodes.size() / 3,
                                                       tris.data(), (int64_t)bt.size(), bt.data(), bv.data(),
                                                       adm.data(), has.data(), SBH_NORMALS_VERTEX_ORDER, out);
        if (st != SBH_OK) throw Error(st, sbh_last_error());
    });
}

}  // extern "C"
