// Scene file readers and writers (libsplat_scene.so): PLY (3DGS layout) and
// .splat records into the activated structure of arrays, and back.
//
// Parsing follows ply::parse (src/ply.rs:107-379) step for step: the header
// is located and validated first, columns are resolved by name (the last
// property of a name wins, as a HashMap built in order keeps it), and every
// value goes through f64 before its f32 conversion. Binary bodies are split
// into contiguous vertex ranges decoded by host threads; ASCII bodies are
// tokenized in one pass. Only the columns the scene uses are decoded from a
// binary row: the others cannot change the result.
//
// f32 arithmetic without contraction (built with -ffp-contract=off); exp and
// ln from the host libm, as Rust's f32::exp / f32::ln on Linux.

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <vector>

#include <sys/mman.h>
#include <sys/stat.h>

#include "splat_io.h"
#include "splat_scene.h"

// Float array whose new elements are left uninitialised: the decoding threads
// write (and first touch) every element, so no serial zero fill precedes them.
class FloatBuf {
public:
    float *data() { return p_.get(); }
    const float *data() const { return p_.get(); }
    float &operator[](uint64_t i) { return p_[i]; }
    uint64_t size() const { return n_; }
    void reserve(uint64_t c) {
        if (c <= cap_) return;
        std::unique_ptr<float[]> q(new float[c]);
        if (n_) std::memcpy(q.get(), p_.get(), n_ * sizeof(float));
        p_ = std::move(q);
        cap_ = c;
    }
    void resize(uint64_t m) {
        if (m > cap_) reserve(std::max(m, 2 * cap_));
        n_ = m;
    }

private:
    std::unique_ptr<float[]> p_;
    uint64_t n_ = 0, cap_ = 0;
};

struct sp_io_scene {
    uint64_t n = 0;
    uint32_t sh_degree = 0;
    uint32_t rest = 0;  // coefficients per gaussian
    FloatBuf positions, scales, rotations, opacities, colors, sh_rest, cov3d;
    bool finalized = false;

    void reserve(uint64_t m) {
        positions.reserve(3 * m);
        scales.reserve(3 * m);
        rotations.reserve(4 * m);
        opacities.reserve(m);
        colors.reserve(3 * m);
        sh_rest.reserve(3 * (uint64_t)rest * m);
    }
    void resize(uint64_t m) {
        n = m;
        positions.resize(3 * m);
        scales.resize(3 * m);
        rotations.resize(4 * m);
        opacities.resize(m);
        colors.resize(3 * m);
        sh_rest.resize(3 * (uint64_t)rest * m);
    }
};

namespace {

constexpr float kShC0 = 0.2820948f;  // sh::SH_C0 (src/sh.rs:5)

thread_local std::string g_err;

int fail(int code, std::string msg) {
    g_err = std::move(msg);
    return code;
}

// ---------------------------------------------------------------------------
// Text helpers: UTF-8 validation and the whitespace classes Rust's str
// methods use (char::is_whitespace for the header, u8::is_ascii_whitespace
// for an ASCII body).

bool utf8_valid(const uint8_t *p, size_t n) {
    size_t i = 0;
    while (i < n) {
        const uint8_t c = p[i];
        if (c < 0x80) {
            ++i;
            continue;
        }
        size_t len;
        uint32_t cp;
        if ((c & 0xE0) == 0xC0) {
            len = 2;
            cp = c & 0x1F;
        } else if ((c & 0xF0) == 0xE0) {
            len = 3;
            cp = c & 0x0F;
        } else if ((c & 0xF8) == 0xF0) {
            len = 4;
            cp = c & 0x07;
        } else {
            return false;
        }
        if (i + len > n) return false;
        for (size_t k = 1; k < len; ++k) {
            if ((p[i + k] & 0xC0) != 0x80) return false;
            cp = (cp << 6) | (p[i + k] & 0x3F);
        }
        // Overlong forms, surrogates and code points above U+10FFFF are invalid.
        if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && cp < 0x10000)) return false;
        if ((cp >= 0xD800 && cp <= 0xDFFF) || cp > 0x10FFFF) return false;
        i += len;
    }
    return true;
}

// Decodes the code point at s[i] of valid UTF-8; returns its length.
size_t utf8_at(std::string_view s, size_t i, uint32_t *cp) {
    const uint8_t c = (uint8_t)s[i];
    if (c < 0x80) {
        *cp = c;
        return 1;
    }
    const size_t len = (c & 0xE0) == 0xC0 ? 2 : (c & 0xF0) == 0xE0 ? 3 : 4;
    uint32_t v = c & (len == 2 ? 0x1F : len == 3 ? 0x0F : 0x07);
    for (size_t k = 1; k < len; ++k) v = (v << 6) | ((uint8_t)s[i + k] & 0x3F);
    *cp = v;
    return len;
}

// Unicode White_Space (char::is_whitespace).
bool unicode_ws(uint32_t c) {
    return (c >= 0x09 && c <= 0x0D) || c == 0x20 || c == 0x85 || c == 0xA0 || c == 0x1680 ||
           (c >= 0x2000 && c <= 0x200A) || c == 0x2028 || c == 0x2029 || c == 0x202F || c == 0x205F ||
           c == 0x3000;
}

// str::split_whitespace.
std::vector<std::string_view> split_ws(std::string_view s) {
    std::vector<std::string_view> out;
    size_t i = 0, start = std::string_view::npos;
    while (i < s.size()) {
        uint32_t cp;
        const size_t len = utf8_at(s, i, &cp);
        if (unicode_ws(cp)) {
            if (start != std::string_view::npos) out.push_back(s.substr(start, i - start));
            start = std::string_view::npos;
        } else if (start == std::string_view::npos) {
            start = i;
        }
        i += len;
    }
    if (start != std::string_view::npos) out.push_back(s.substr(start));
    return out;
}

// u8::is_ascii_whitespace: space, tab, LF, FF, CR (not VT).
inline bool ascii_ws(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\f' || c == '\r'; }

bool ieq(std::string_view a, const char *b) {
    const size_t n = std::strlen(b);
    if (a.size() != n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)a[i]) != b[i]) return false;
    return true;
}

// <f64 as FromStr>::from_str: [+-] then inf / infinity / nan (any case) or a
// decimal with at least one digit and an optional exponent. Correctly rounded
// (strtod on the validated token).
bool parse_f64(std::string_view t, double *out) {
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) {
        neg = t[i] == '-';
        ++i;
    }
    const std::string_view r = t.substr(i);
    if (ieq(r, "inf") || ieq(r, "infinity")) {
        *out = neg ? -INFINITY : INFINITY;
        return true;
    }
    if (ieq(r, "nan")) {
        *out = neg ? -NAN : NAN;
        return true;
    }
    size_t digits = 0;
    while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i, ++digits;
    if (i < t.size() && t[i] == '.') {
        ++i;
        while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i, ++digits;
    }
    if (digits == 0) return false;
    if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
        ++i;
        if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
        size_t ed = 0;
        while (i < t.size() && t[i] >= '0' && t[i] <= '9') ++i, ++ed;
        if (ed == 0) return false;
    }
    if (i != t.size()) return false;
    const std::string tok(t);
    *out = std::strtod(tok.c_str(), nullptr);
    return true;
}

// <usize as FromStr>::from_str: an optional '+', then decimal digits, no overflow.
bool parse_count(std::string_view t, uint64_t *out) {
    size_t i = 0;
    if (i < t.size() && t[i] == '+') ++i;
    if (i == t.size()) return false;
    uint64_t v = 0;
    for (; i < t.size(); ++i) {
        if (t[i] < '0' || t[i] > '9') return false;
        const uint64_t d = (uint64_t)(t[i] - '0');
        if (v > (UINT64_MAX - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

// ---------------------------------------------------------------------------
// PLY header (src/ply.rs:18-185).

enum class Fmt { Ascii, Le, Be };
enum class Ty : uint8_t { I8, U8, I16, U16, I32, U32, F32, F64 };

bool parse_type(std::string_view s, Ty *t) {
    if (s == "char" || s == "int8") *t = Ty::I8;
    else if (s == "uchar" || s == "uint8") *t = Ty::U8;
    else if (s == "short" || s == "int16") *t = Ty::I16;
    else if (s == "ushort" || s == "uint16") *t = Ty::U16;
    else if (s == "int" || s == "int32") *t = Ty::I32;
    else if (s == "uint" || s == "uint32") *t = Ty::U32;
    else if (s == "float" || s == "float32") *t = Ty::F32;
    else if (s == "double" || s == "float64") *t = Ty::F64;
    else return false;
    return true;
}

uint32_t ty_size(Ty t) {
    switch (t) {
    case Ty::I8:
    case Ty::U8: return 1;
    case Ty::I16:
    case Ty::U16: return 2;
    case Ty::F64: return 8;
    default: return 4;
    }
}

// ScalarType::read: the value as f64.
inline double read_scalar(Ty t, const uint8_t *b, bool le) {
    uint8_t v[8];
    const uint32_t n = ty_size(t);
    if (le) {
        std::memcpy(v, b, n);
    } else {
        for (uint32_t k = 0; k < n; ++k) v[k] = b[n - 1 - k];
    }
    switch (t) {
    case Ty::I8: return (double)(int8_t)v[0];
    case Ty::U8: return (double)v[0];
    case Ty::I16: { int16_t x; std::memcpy(&x, v, 2); return (double)x; }
    case Ty::U16: { uint16_t x; std::memcpy(&x, v, 2); return (double)x; }
    case Ty::I32: { int32_t x; std::memcpy(&x, v, 4); return (double)x; }
    case Ty::U32: { uint32_t x; std::memcpy(&x, v, 4); return (double)x; }
    case Ty::F32: { float x; std::memcpy(&x, v, 4); return (double)x; }
    default: { double x; std::memcpy(&x, v, 8); return x; }
    }
}

struct Prop {
    std::string name;
    Ty ty;
    bool list;
};

struct Elem {
    std::string name;
    uint64_t count;
    std::vector<Prop> props;
};

struct Header {
    Fmt fmt;
    std::vector<Elem> elems;
    uint64_t body;
};

int parse_header(const uint8_t *data, uint64_t len, Header *h) {
    static const char kEnd[] = "end_header";
    const uint64_t m = sizeof(kEnd) - 1;
    // The marker must start within the first 2^20 bytes.
    uint64_t pos = UINT64_MAX;
    if (len >= m) {
        const uint64_t last = std::min<uint64_t>(len - m, (1u << 20) - 1);
        for (uint64_t i = 0; i <= last; ++i) {
            if (data[i] == 'e' && std::memcmp(data + i, kEnd, m) == 0) {
                pos = i;
                break;
            }
        }
    }
    if (pos == UINT64_MAX) return fail(SP_E_FORMAT, "PLY header missing end_header");
    // The body starts after the newline that ends the end_header line.
    uint64_t body = pos + m;
    while (body < len && data[body] != '\n') ++body;
    h->body = body + 1;
    if (!utf8_valid(data, pos)) return fail(SP_E_FORMAT, "PLY header is not UTF-8");
    const std::string_view text((const char *)data, pos);

    bool have_fmt = false, first = true;
    size_t at = 0;
    while (at < text.size()) {
        size_t nl = text.find('\n', at);
        if (nl == std::string_view::npos) nl = text.size();
        const std::vector<std::string_view> tok = split_ws(text.substr(at, nl - at));
        at = nl + 1;
        if (first) {
            first = false;
            if (tok.size() != 1 || tok[0] != "ply") return fail(SP_E_FORMAT, "not a PLY file (missing 'ply' magic)");
            continue;
        }
        if (tok.empty() || tok[0] == "comment" || tok[0] == "obj_info") continue;
        if (tok[0] == "format") {
            const std::string_view f = tok.size() > 1 ? tok[1] : std::string_view();
            if (f == "ascii") h->fmt = Fmt::Ascii;
            else if (f == "binary_little_endian") h->fmt = Fmt::Le;
            else if (f == "binary_big_endian") h->fmt = Fmt::Be;
            else return fail(SP_E_FORMAT, "unsupported PLY format '" + std::string(f) + "'");
            have_fmt = true;
        } else if (tok[0] == "element") {
            if (tok.size() < 2) return fail(SP_E_FORMAT, "element without name");
            if (tok.size() < 3) return fail(SP_E_FORMAT, "element without count");
            uint64_t count;
            if (!parse_count(tok[2], &count)) return fail(SP_E_FORMAT, "bad element count");
            h->elems.push_back({std::string(tok[1]), count, {}});
        } else if (tok[0] == "property") {
            if (h->elems.empty()) return fail(SP_E_FORMAT, "property before element");
            if (tok.size() < 2) return fail(SP_E_FORMAT, "property without type");
            Prop p;
            if (tok[1] == "list") {
                Ty ct;
                if (!parse_type(tok.size() > 2 ? tok[2] : "", &ct) || !parse_type(tok.size() > 3 ? tok[3] : "", &p.ty))
                    return fail(SP_E_FORMAT, "unsupported PLY property type");
                if (tok.size() < 5) return fail(SP_E_FORMAT, "list without name");
                p.name = std::string(tok[4]);
                p.list = true;
            } else {
                if (tok.size() < 3) return fail(SP_E_FORMAT, "property without name");
                if (!parse_type(tok[1], &p.ty))
                    return fail(SP_E_FORMAT, "unsupported PLY property type '" + std::string(tok[1]) + "'");
                p.name = std::string(tok[2]);
                p.list = false;
            }
            h->elems.back().props.push_back(std::move(p));
        } else {
            return fail(SP_E_FORMAT, "unexpected PLY header keyword '" + std::string(tok[0]) + "'");
        }
    }
    if (first) return fail(SP_E_FORMAT, "not a PLY file (missing 'ply' magic)");
    if (!have_fmt) return fail(SP_E_FORMAT, "PLY header has no format line");
    return SP_OK;
}

// Columns (src/ply.rs:187-229): property index of each understood field, -1 if absent.
struct Columns {
    int pos[3];
    int dc[3];
    int rgb[3];
    int opacity;
    int scale[3];
    int rot[4];
    std::vector<int> rest;
};

int resolve_columns(const std::vector<Prop> &props, Columns *c) {
    std::unordered_map<std::string, int> idx;
    for (size_t i = 0; i < props.size(); ++i) idx[props[i].name] = (int)i;  // the last one of a name wins
    const auto get = [&](const std::string &n) {
        auto it = idx.find(n);
        return it == idx.end() ? -1 : it->second;
    };
    const auto all = [&](int *dst, std::initializer_list<const char *> names) {
        int k = 0;
        bool ok = true;
        for (const char *n : names) ok = (dst[k++] = get(n)) >= 0 && ok;
        if (!ok)
            for (int j = 0; j < k; ++j) dst[j] = -1;
    };
    c->rest.clear();
    for (;;) {
        const int i = get("f_rest_" + std::to_string(c->rest.size()));
        if (i < 0) break;
        c->rest.push_back(i);
    }
    for (const Prop &p : props)
        if (p.name.compare(0, 7, "packed_") == 0)
            return fail(SP_E_FORMAT,
                        "compressed PLY (SuperSplat packed format) is not supported; decompress it first");
    all(c->rot, {"rot_0", "rot_1", "rot_2", "rot_3"});
    const char *xyz[3] = {"x", "y", "z"};
    for (int k = 0; k < 3; ++k) {
        c->pos[k] = get(xyz[k]);
        if (c->pos[k] < 0) return fail(SP_E_FORMAT, std::string("PLY vertex missing '") + xyz[k] + "'");
    }
    all(c->dc, {"f_dc_0", "f_dc_1", "f_dc_2"});
    all(c->rgb, {"red", "green", "blue"});
    c->opacity = get("opacity");
    all(c->scale, {"scale_0", "scale_1", "scale_2"});
    return SP_OK;
}

// sh::degree_from_rest_count (src/sh.rs:30-32), 0 when no degree matches.
uint32_t degree_of_rest(size_t rest) {
    for (uint32_t d = 0; d <= 3; ++d)
        if (3 * ((size_t)(d + 1) * (d + 1) - 1) == rest) return d;
    return 0;
}

inline float sigmoid(float x) { return 1.0f / (1.0f + std::exp(-x)); }

// The per-vertex mapping of columns to the scene (the `emit` closure,
// src/ply.rs:275-315), for a row readable as f64 by column index.
struct Emitter {
    const Columns &c;
    float rgb_scale;
    uint32_t pc;  // rest coefficients per channel

    template <class Row>
    void operator()(const Row &row, sp_io_scene &s, uint64_t i) const {
        const auto f = [&](int col) { return (float)row(col); };
        float *pos = &s.positions[3 * i];
        for (int k = 0; k < 3; ++k) pos[k] = f(c.pos[k]);
        float *col = &s.colors[3 * i];
        if (c.dc[0] >= 0) {
            for (int k = 0; k < 3; ++k) col[k] = f(c.dc[k]) * kShC0 + 0.5f;
        } else if (c.rgb[0] >= 0) {
            for (int k = 0; k < 3; ++k) col[k] = f(c.rgb[k]) * rgb_scale;
        } else {
            col[0] = col[1] = col[2] = 0.8f;
        }
        s.opacities[i] = c.opacity >= 0 ? sigmoid(f(c.opacity)) : 1.0f;
        float *sc = &s.scales[3 * i];
        for (int k = 0; k < 3; ++k) sc[k] = c.scale[0] >= 0 ? std::exp(f(c.scale[k])) : 0.01f;
        float *q = &s.rotations[4 * i];
        if (c.rot[0] >= 0) {
            // 3DGS stores quaternions as (w, x, y, z).
            q[0] = f(c.rot[1]);
            q[1] = f(c.rot[2]);
            q[2] = f(c.rot[3]);
            q[3] = f(c.rot[0]);
        } else {
            q[0] = q[1] = q[2] = 0.0f;
            q[3] = 1.0f;
        }
        // f_rest is channel-major (all R, then G, then B); the scene interleaves RGB per coefficient.
        float *r = s.sh_rest.data() + 3 * (uint64_t)pc * i;
        for (uint32_t k = 0; k < pc; ++k)
            for (int ch = 0; ch < 3; ++ch) r[3 * k + ch] = f(c.rest[ch * pc + k]);
    }
};

uint32_t thread_count(uint32_t threads, uint64_t work) {
    uint32_t t = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    // Below ~16k rows per thread the start-up cost dominates.
    const uint64_t cap = std::max<uint64_t>(1, work / 16384);
    return (uint32_t)std::min<uint64_t>(t, cap);
}

template <class Fn>
void parallel_ranges(uint64_t n, uint32_t threads, Fn fn) {
    const uint32_t t = thread_count(threads, n);
    if (t <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t step = (n + t - 1) / t;
    for (uint32_t k = 0; k < t; ++k) {
        const uint64_t a = k * step, b = std::min(n, a + step);
        if (a >= b) break;
        pool.emplace_back([=] { fn(a, b); });
    }
    for (auto &th : pool) th.join();
}

int ply_parse(const uint8_t *data, uint64_t len, uint32_t threads, sp_io_scene *s) {
    Header h;
    if (int e = parse_header(data, len, &h)) return e;
    size_t vi = h.elems.size();
    for (size_t i = 0; i < h.elems.size(); ++i)
        if (h.elems[i].name == "vertex") {
            vi = i;
            break;
        }
    if (vi == h.elems.size()) return fail(SP_E_FORMAT, "PLY has no vertex element");
    // Elements stored before `vertex` are skipped; only fixed-size ones can be.
    uint64_t offset = h.body;
    for (size_t i = 0; i < vi; ++i) {
        const Elem &el = h.elems[i];
        bool has_list = false;
        uint64_t stride = 0;
        for (const Prop &p : el.props) {
            has_list |= p.list;
            stride += ty_size(p.ty);
        }
        if (h.fmt == Fmt::Ascii || has_list)
            return fail(SP_E_FORMAT, "PLY element '" + el.name + "' before vertex is not supported");
        if (stride && el.count > (UINT64_MAX - offset) / stride) return fail(SP_E_FORMAT, "PLY too large");
        offset += stride * el.count;
    }
    const Elem &el = h.elems[vi];
    Columns cols;
    if (int e = resolve_columns(el.props, &cols)) return e;
    const uint64_t n = el.count;
    s->sh_degree = degree_of_rest(cols.rest.size());
    s->rest = (s->sh_degree + 1) * (s->sh_degree + 1) - 1;
    const float rgb_scale = cols.rgb[0] >= 0 && el.props[cols.rgb[0]].ty == Ty::U8 ? 1.0f / 255.0f : 1.0f;
    const Emitter emit{cols, rgb_scale, s->rest};

    if (h.fmt == Fmt::Ascii) {
        const uint64_t start = std::min(offset, len);
        if (!utf8_valid(data + start, len - start)) return fail(SP_E_FORMAT, "ASCII PLY body is not UTF-8");
        const uint8_t *p = data + start, *end = data + len;
        const auto next = [&](std::string_view *tok) {
            while (p < end && ascii_ws(*p)) ++p;
            if (p == end) return false;
            const uint8_t *b = p;
            while (p < end && !ascii_ws(*p)) ++p;
            *tok = std::string_view((const char *)b, (size_t)(p - b));
            return true;
        };
        std::vector<double> row(el.props.size());
        // Never trust the header count for allocation: a vertex takes at least
        // one byte per property.
        s->reserve(std::min<uint64_t>(n, (len - start) / std::max<size_t>(1, el.props.size()) + 1));
        for (uint64_t v = 0; v < n; ++v) {
            for (size_t pi = 0; pi < el.props.size(); ++pi) {
                std::string_view tok;
                double x;
                if (!next(&tok)) return fail(SP_E_FORMAT, "ASCII PLY truncated at vertex " + std::to_string(v));
                if (!parse_f64(tok, &x)) return fail(SP_E_FORMAT, "bad number in ASCII PLY");
                if (el.props[pi].list) {
                    // `as usize`: saturating, NaN and negatives to 0.
                    const uint64_t cnt = !(x > 0) ? 0 : x >= 18446744073709551616.0 ? UINT64_MAX : (uint64_t)x;
                    for (uint64_t k = 0; k < cnt; ++k) {
                        if (!next(&tok)) return fail(SP_E_FORMAT, "ASCII PLY truncated at vertex " + std::to_string(v));
                        if (!parse_f64(tok, &x)) return fail(SP_E_FORMAT, "bad number in ASCII PLY");
                    }
                    row[pi] = 0.0;
                } else {
                    row[pi] = x;
                }
            }
            s->resize(v + 1);
            emit([&](int col) { return row[col]; }, *s, v);
        }
        s->n = n;
        return SP_OK;
    }

    for (const Prop &p : el.props)
        if (p.list) return fail(SP_E_FORMAT, "list properties on vertices are not supported");
    std::vector<uint32_t> offs(el.props.size());
    uint64_t stride = 0;
    for (size_t i = 0; i < el.props.size(); ++i) {
        offs[i] = (uint32_t)stride;
        stride += ty_size(el.props[i].ty);
    }
    if (stride && n > (UINT64_MAX - offset) / stride) return fail(SP_E_FORMAT, "PLY vertex count overflows");
    const uint64_t need = offset + stride * n;
    if (len < need)
        return fail(SP_E_FORMAT, "PLY body truncated: need " + std::to_string(need) + " bytes, file has " +
                                     std::to_string(len));
    s->resize(n);
    const bool le = h.fmt == Fmt::Le;
    const uint8_t *body = data + offset;
    parallel_ranges(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t v = a; v < b; ++v) {
            const uint8_t *rec = body + v * stride;
            emit([&](int col) { return read_scalar(el.props[col].ty, rec + offs[col], le); }, *s, v);
        }
    });
    return SP_OK;
}

// splat_format::parse (src/splat_format.rs:18-44).
int splat_parse(const uint8_t *data, uint64_t len, uint32_t threads, sp_io_scene *s) {
    if (len % 32 != 0)
        return fail(SP_E_FORMAT, ".splat file size " + std::to_string(len) + " is not a multiple of 32 bytes");
    const uint64_t n = len / 32;
    s->sh_degree = 0;
    s->rest = 0;
    s->resize(n);
    parallel_ranges(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) {
            const uint8_t *r = data + 32 * i;
            float f[6];
            std::memcpy(f, r, 24);  // little-endian host
            for (int k = 0; k < 3; ++k) {
                s->positions[3 * i + k] = f[k];
                s->scales[3 * i + k] = std::fabs(f[3 + k]);
                s->colors[3 * i + k] = (float)r[24 + k] / 255.0f;
            }
            s->opacities[i] = (float)r[27] / 255.0f;
            const auto q = [](uint8_t b) { return ((float)b - 128.0f) / 128.0f; };
            s->rotations[4 * i] = q(r[29]);
            s->rotations[4 * i + 1] = q(r[30]);
            s->rotations[4 * i + 2] = q(r[31]);
            s->rotations[4 * i + 3] = q(r[28]);
        }
    });
    return SP_OK;
}

bool arrays_ok(const sp_io_arrays *a) {
    if (!a || (a->n && (!a->positions || !a->scales || !a->rotations || !a->opacities || !a->colors))) return false;
    const uint32_t rest = (a->sh_degree + 1) * (a->sh_degree + 1) - 1;
    return a->sh_degree <= 3 && a->rest == rest && (rest == 0 || a->n == 0 || a->sh_rest);
}

inline void put_f32(uint8_t *&o, float v) {
    std::memcpy(o, &v, 4);  // little-endian host
    o += 4;
}

// scene::logit (src/scene.rs:188-191).
inline float logit(float p) {
    if (p < 1e-6f) p = 1e-6f;
    if (p > 1.0f - 1e-6f) p = 1.0f - 1e-6f;
    return std::log(p / (1.0f - p));
}

}  // namespace

extern "C" {

int sp_ply_parse(const uint8_t *data, uint64_t len, uint32_t threads, sp_io_scene **out) {
    if (!out || (!data && len)) return fail(SP_E_INVALID, "null argument");
    *out = nullptr;
    auto *s = new sp_io_scene;
    if (int e = ply_parse(data, len, threads, s)) {
        delete s;
        return e;
    }
    g_err.clear();
    *out = s;
    return SP_OK;
}

int sp_splat_parse(const uint8_t *data, uint64_t len, uint32_t threads, sp_io_scene **out) {
    if (!out || (!data && len)) return fail(SP_E_INVALID, "null argument");
    *out = nullptr;
    auto *s = new sp_io_scene;
    if (int e = splat_parse(data, len, threads, s)) {
        delete s;
        return e;
    }
    g_err.clear();
    *out = s;
    return SP_OK;
}

int sp_io_finalize(sp_io_scene *s) {
    if (!s) return fail(SP_E_INVALID, "null scene");
    s->cov3d.resize(6 * s->n);
    // Per-gaussian and independent: contiguous ranges on host threads.
    parallel_ranges(s->n, 0, [&](uint64_t a, uint64_t b) {
        sp_finalize(b - a, s->scales.data() + 3 * a, s->rotations.data() + 4 * a, s->opacities.data() + a,
                    s->cov3d.data() + 6 * a);
    });
    s->finalized = true;
    return SP_OK;
}

int sp_scene_load(const char *path, uint32_t threads, sp_io_scene **out) {
    if (!path || !out) return fail(SP_E_INVALID, "null argument");
    *out = nullptr;
    // Path::extension: after the last '.' of the file name, unless that dot leads the name.
    const char *name = std::strrchr(path, '/');
    name = name ? name + 1 : path;
    const char *dot = std::strrchr(name, '.');
    std::string ext = dot && dot != name ? std::string(dot + 1) : std::string();
    for (char &ch : ext) ch = (char)std::tolower((unsigned char)ch);

    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SP_E_IO, std::string("reading ") + path + ": " + std::strerror(errno));
    // A regular file is mapped, so the decoding threads fault its pages in
    // parallel; anything else is read into a growing buffer.
    struct stat st;
    const uint8_t *bytes = nullptr;
    size_t len = 0;
    void *map = MAP_FAILED;
    std::unique_ptr<uint8_t[]> data;
    if (fstat(fileno(f), &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
        len = (size_t)st.st_size;
        map = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fileno(f), 0);
        if (map != MAP_FAILED) bytes = (const uint8_t *)map;
    }
    if (!bytes) {
        size_t cap = (size_t)1 << 20;
        len = 0;
        data.reset(new uint8_t[cap]);
        for (;;) {
            if (len == cap) {
                std::unique_ptr<uint8_t[]> bigger(new uint8_t[2 * cap]);
                std::memcpy(bigger.get(), data.get(), len);
                data = std::move(bigger);
                cap *= 2;
            }
            const size_t got = std::fread(data.get() + len, 1, cap - len, f);
            if (got == 0) break;
            len += got;
        }
        bytes = data.get();
    }
    const bool err = std::ferror(f);
    std::fclose(f);
    struct Unmap {
        void *p;
        size_t n;
        ~Unmap() {
            if (p != MAP_FAILED) munmap(p, n);
        }
    } unmap{map, len};
    if (err) return fail(SP_E_IO, std::string("reading ") + path);

    const bool ply_magic = len >= 3 && std::memcmp(bytes, "ply", 3) == 0;
    sp_io_scene *s = nullptr;
    int e;
    if (ext == "splat") e = sp_splat_parse(bytes, len, threads, &s);
    else if (ext == "ply" || ply_magic) e = sp_ply_parse(bytes, len, threads, &s);
    else return fail(SP_E_FORMAT, std::string("unknown file type (expected .ply or .splat): ") + path);
    if (e) return e;
    sp_io_finalize(s);
    if (s->n == 0) {
        delete s;
        return fail(SP_E_FORMAT, std::string(path) + " contains no gaussians");
    }
    g_err.clear();
    *out = s;
    return SP_OK;
}

void sp_io_arrays_of(sp_io_scene *s, sp_io_arrays *out) {
    if (!out) return;
    std::memset(out, 0, sizeof *out);
    if (!s) return;
    out->n = s->n;
    out->sh_degree = s->sh_degree;
    out->rest = s->rest;
    out->positions = s->positions.data();
    out->scales = s->scales.data();
    out->rotations = s->rotations.data();
    out->opacities = s->opacities.data();
    out->colors = s->colors.data();
    out->sh_rest = s->rest ? s->sh_rest.data() : nullptr;
    out->cov3d = s->finalized ? s->cov3d.data() : nullptr;
}

void sp_io_free(sp_io_scene *s) { delete s; }

const char *sp_io_last_error(void) { return g_err.c_str(); }

// ply::write (src/ply.rs:382-431).
int sp_ply_write(const sp_io_arrays *a, uint8_t *out, uint64_t cap, uint64_t *written) {
    if (!arrays_ok(a) || !written) return fail(SP_E_INVALID, "inconsistent scene arrays");
    const uint32_t pc = a->rest;
    std::string h = "ply\nformat binary_little_endian 1.0\ncomment generated by splat-tui\n";
    h += "element vertex " + std::to_string(a->n) + "\n";
    for (const char *p : {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"})
        h += std::string("property float ") + p + "\n";
    for (uint32_t i = 0; i < 3 * pc; ++i) h += "property float f_rest_" + std::to_string(i) + "\n";
    for (const char *p : {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"})
        h += std::string("property float ") + p + "\n";
    h += "end_header\n";
    const uint64_t row = (17 + 3 * (uint64_t)pc) * 4;
    *written = h.size() + row * a->n;
    if (cap < *written) return fail(SP_E_CAPACITY, "output buffer too small");
    if (!out) return fail(SP_E_INVALID, "null output");
    std::memcpy(out, h.data(), h.size());
    uint8_t *body = out + h.size();
    parallel_ranges(a->n, 0, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            uint8_t *o = body + i * row;
            const float *p = a->positions + 3 * i, *c = a->colors + 3 * i, *s = a->scales + 3 * i;
            const float *q = a->rotations + 4 * i;
            for (int k = 0; k < 3; ++k) put_f32(o, p[k]);
            for (int k = 0; k < 3; ++k) put_f32(o, 0.0f);
            for (int k = 0; k < 3; ++k) put_f32(o, (c[k] - 0.5f) / kShC0);
            const float *r = pc ? a->sh_rest + 3 * (uint64_t)pc * i : nullptr;
            for (int ch = 0; ch < 3; ++ch)
                for (uint32_t k = 0; k < pc; ++k) put_f32(o, r[3 * k + ch]);
            put_f32(o, logit(a->opacities[i]));
            for (int k = 0; k < 3; ++k) put_f32(o, std::log(std::fmax(s[k], 1e-12f)));
            put_f32(o, q[3]);
            for (int k = 0; k < 3; ++k) put_f32(o, q[k]);
        }
    });
    g_err.clear();
    return SP_OK;
}

// splat_format::write (src/splat_format.rs:46-62).
int sp_splat_write(const sp_io_arrays *a, uint8_t *out, uint64_t cap, uint64_t *written) {
    if (!arrays_ok(a) || !written) return fail(SP_E_INVALID, "inconsistent scene arrays");
    *written = 32 * a->n;
    if (cap < *written) return fail(SP_E_CAPACITY, "output buffer too small");
    if (!out && a->n) return fail(SP_E_INVALID, "null output");
    // (v * k + b).round().clamp(0, 255) as u8: NaN stays NaN through clamp and casts to 0.
    const auto quant = [](float v) -> uint8_t {
        float r = std::round(v);
        if (r < 0.0f) r = 0.0f;
        if (r > 255.0f) r = 255.0f;
        return r == r ? (uint8_t)r : 0;
    };
    parallel_ranges(a->n, 0, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            uint8_t *o = out + 32 * i;
            for (int k = 0; k < 3; ++k) put_f32(o, a->positions[3 * i + k]);
            for (int k = 0; k < 3; ++k) put_f32(o, a->scales[3 * i + k]);
            for (int k = 0; k < 3; ++k) *o++ = quant(a->colors[3 * i + k] * 255.0f);
            *o++ = quant(a->opacities[i] * 255.0f);
            // Quat::normalize with glam's SSE2 4-lane length ((x^2 + z^2) + (y^2 + w^2)), times 1/length.
            const float *q = a->rotations + 4 * i;
            const float l = std::sqrt((q[0] * q[0] + q[2] * q[2]) + (q[1] * q[1] + q[3] * q[3]));
            const float rl = 1.0f / l;
            const float u[4] = {q[0] * rl, q[1] * rl, q[2] * rl, q[3] * rl};
            *o++ = quant(u[3] * 128.0f + 128.0f);
            for (int k = 0; k < 3; ++k) *o++ = quant(u[k] * 128.0f + 128.0f);
        }
    });
    g_err.clear();
    return SP_OK;
}

}  // extern "C"
