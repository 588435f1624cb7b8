// capi_state.cpp -- state interchange: LatticeStateDefault<D> and
// LatticeStateEFSyncDefault<LatticeStateDefault<D>, D> in the serde data
// model the reference derives (src/simulation/state.rs:653-659,1046-1062,
// src/field.rs:26-31,582-586,1022-1027, src/lattice.rs:42-49), written and
// read in two serde formats:
//
//   LQCD_FORMAT_BINCODE  bincode 1.x `serialize` (fixed-width little-endian
//                        integers, u64 lengths before sequences, tuples and
//                        structs as their fields in order);
//   LQCD_FORMAT_JSON     serde_json (structs as objects, tuples as arrays).
//
// The data model, from the derives and the serde impls of the crates the
// reference pins (nalgebra 0.31 `serde-serialize`: a matrix is the tuple of
// its R*C entries in column-major order; num-complex 0.4: a complex is the
// tuple (re, im)):
//
//   LatticeStateDefault  { lattice: { size: f64, dim: usize }, beta: f64,
//                          link_matrix: { data: Vec<[(re, im); 9]> } }
//   EField               { data: Vec<[ { data: [f64; 8] } ; D]> }
//   LatticeStateEFSyncDefault { e_field: EField, t: usize,
//                               lattice_state: LatticeStateDefault }
//
// D is a const generic and is not serialized: the reader is told D and checks
// the lengths against dim^D.  Links and E move through the host layouts of
// lqcd_links_download / lqcd_efield_download, which are the same orders.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "capi_util.hpp"
#include "internal.hpp"

using namespace lqcd;

namespace {

struct File {
    FILE* f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int io_fail(const std::string& what, const char* path)
{
    return fail(LQCD_ERR_IO, what + " (" + path + ")");
}

// no C++ exception crosses the C ABI: an allocation failure becomes
// LQCD_ERR_OUT_OF_MEMORY, anything else thrown LQCD_ERR_IO
template <typename Fn>
int guarded(Fn&& fn)
{
    try {
        return fn();
    } catch (const std::bad_alloc&) {
        return fail(LQCD_ERR_OUT_OF_MEMORY, "host allocation failed while reading or writing a state");
    } catch (const std::exception& e) {
        return fail(LQCD_ERR_IO, std::string("state I/O: ") + e.what());
    }
}

// ------------------------------------------------------------------ writers

class Writer {
public:
    explicit Writer(FILE* f) : f_(f) { buf_.reserve(1 << 20); }
    ~Writer() { flush(); }
    void bytes(const void* p, size_t n)
    {
        const char* c = static_cast<const char*>(p);
        buf_.insert(buf_.end(), c, c + n);
        if (buf_.size() >= (1u << 20)) flush();
    }
    void text(const char* s) { bytes(s, std::strlen(s)); }
    void u64(uint64_t v)
    {
        unsigned char b[8];
        for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
        bytes(b, 8);
    }
    void f64(double v)
    {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        u64(u);
    }
    // serde_json writes the shortest representation that reads back to the
    // same double, with a ".0" on integral values, and null for NaN / inf
    void json_f64(double v)
    {
        if (!std::isfinite(v)) {
            text("null");
            return;
        }
        char s[40];
        auto r = std::to_chars(s, s + sizeof(s) - 3, v);
        *r.ptr = '\0';
        bool plain = true;
        for (char* c = s; c < r.ptr; ++c)
            if (*c == '.' || *c == 'e' || *c == 'E') plain = false;
        if (plain) std::strcat(s, ".0");
        text(s);
    }
    void json_u64(uint64_t v)
    {
        char s[24];
        auto r = std::to_chars(s, s + sizeof(s) - 1, v);
        *r.ptr = '\0';
        text(s);
    }
    bool ok() { return flush() && !std::ferror(f_); }

private:
    bool flush()
    {
        if (buf_.empty()) return true;
        const size_t n = std::fwrite(buf_.data(), 1, buf_.size(), f_);
        const bool good = n == buf_.size();
        buf_.clear();
        return good;
    }
    FILE* f_;
    std::vector<char> buf_;
};

struct HostState {
    int dim = 0;
    double size = 0.0, beta = 0.0;
    uint64_t points = 0, t = 0;
    std::vector<double> links;   // 18 per link
    std::vector<double> efield;  // 8 per link, empty for LatticeStateDefault
};

void write_state_bincode(Writer& w, const HostState& s)
{
    w.f64(s.size);
    w.u64(s.points);
    w.f64(s.beta);
    w.u64(s.links.size() / 18);
    w.bytes(s.links.data(), s.links.size() * sizeof(double));   // little-endian host
}

void write_state_json(Writer& w, const HostState& s)
{
    w.text("{\"lattice\":{\"size\":");
    w.json_f64(s.size);
    w.text(",\"dim\":");
    w.json_u64(s.points);
    w.text("},\"beta\":");
    w.json_f64(s.beta);
    w.text(",\"link_matrix\":{\"data\":[");
    const size_t n = s.links.size() / 18;
    for (size_t l = 0; l < n; ++l) {
        w.text(l ? ",[" : "[");
        for (int k = 0; k < 9; ++k) {
            w.text(k ? ",[" : "[");
            w.json_f64(s.links[18 * l + 2 * k]);
            w.text(",");
            w.json_f64(s.links[18 * l + 2 * k + 1]);
            w.text("]");
        }
        w.text("]");
    }
    w.text("]}}");
}

void write_efield_bincode(Writer& w, const HostState& s)
{
    w.u64(s.efield.size() / (8 * (size_t)s.dim));
    w.bytes(s.efield.data(), s.efield.size() * sizeof(double));
}

void write_efield_json(Writer& w, const HostState& s)
{
    w.text("{\"data\":[");
    const size_t nsite = s.efield.size() / (8 * (size_t)s.dim);
    for (size_t x = 0; x < nsite; ++x) {
        w.text(x ? ",[" : "[");
        for (int mu = 0; mu < s.dim; ++mu) {
            w.text(mu ? ",{\"data\":[" : "{\"data\":[");
            for (int a = 0; a < 8; ++a) {
                if (a) w.text(",");
                w.json_f64(s.efield[(x * s.dim + mu) * 8 + a]);
            }
            w.text("]}");
        }
        w.text("]");
    }
    w.text("]}");
}

// ------------------------------------------------------------------ readers

class Reader {
public:
    explicit Reader(FILE* f) : f_(f) {}
    // bytes not consumed yet of a file of `size` bytes
    uint64_t remaining(uint64_t size) const
    {
        const uint64_t used = read_ - (len_ - pos_);
        return used <= size ? size - used : 0;
    }
    int get()
    {
        if (pos_ == len_) {
            len_ = std::fread(buf_, 1, sizeof(buf_), f_);
            read_ += len_;
            pos_ = 0;
            if (len_ == 0) return EOF;
        }
        return (unsigned char)buf_[pos_++];
    }
    int peek()
    {
        const int c = get();
        if (c != EOF) --pos_;
        return c;
    }
    bool bytes(void* p, size_t n)
    {
        char* d = static_cast<char*>(p);
        while (n) {
            if (pos_ == len_) {
                len_ = std::fread(buf_, 1, sizeof(buf_), f_);
                read_ += len_;
                pos_ = 0;
                if (len_ == 0) return false;
            }
            const size_t k = std::min(n, len_ - pos_);
            std::memcpy(d, buf_ + pos_, k);
            pos_ += k;
            d += k;
            n -= k;
        }
        return true;
    }
    bool u64(uint64_t& v)
    {
        unsigned char b[8];
        if (!bytes(b, 8)) return false;
        v = 0;
        for (int i = 0; i < 8; ++i) v |= (uint64_t)b[i] << (8 * i);
        return true;
    }
    bool f64(double& v)
    {
        uint64_t u;
        if (!u64(u)) return false;
        std::memcpy(&v, &u, 8);
        return true;
    }
    bool at_end() { return peek() == EOF; }

    // ---- JSON tokens
    void ws()
    {
        int c;
        while ((c = peek()) == ' ' || c == '\n' || c == '\r' || c == '\t') get();
    }
    bool expect(char ch)
    {
        ws();
        return get() == ch;
    }
    // ',' or the closing bracket; returns 1 for ',', 0 for close, -1 on error
    int sep(char close)
    {
        ws();
        const int c = get();
        if (c == ',') return 1;
        if (c == close) return 0;
        return -1;
    }
    bool string(std::string& s)
    {
        ws();
        if (get() != '"') return false;
        s.clear();
        for (;;) {
            int c = get();
            if (c == EOF) return false;
            if (c == '"') return true;
            if (c == '\\') {
                c = get();
                if (c == EOF) return false;
            }
            s.push_back((char)c);
        }
    }
    // a JSON number (null is NaN, as serde_json writes NaN / inf)
    bool number(double& v)
    {
        ws();
        char tok[64];
        size_t n = 0;
        int c;
        while ((c = peek()) != EOF && (std::isdigit(c) || c == '-' || c == '+' || c == '.' || c == 'e' ||
                                       c == 'E' || c == 'n' || c == 'u' || c == 'l')) {
            if (n + 1 >= sizeof(tok)) return false;
            tok[n++] = (char)get();
        }
        tok[n] = '\0';
        if (n == 4 && std::strcmp(tok, "null") == 0) {
            v = std::nan("");
            return true;
        }
        auto r = std::from_chars(tok, tok + n, v);
        return r.ec == std::errc() && r.ptr == tok + n;
    }
    bool integer(uint64_t& v)
    {
        double d;
        if (!number(d) || !(d >= 0) || d != std::floor(d) || d > 9.007199254740992e15) return false;
        v = (uint64_t)d;
        return true;
    }
    // skip any JSON value (unknown keys)
    bool skip()
    {
        ws();
        const int c = peek();
        if (c == '{' || c == '[') {
            const char close = c == '{' ? '}' : ']';
            get();
            ws();
            if (peek() == close) {
                get();
                return true;
            }
            for (;;) {
                if (c == '{') {
                    std::string k;
                    if (!string(k) || !expect(':')) return false;
                }
                if (!skip()) return false;
                const int s = sep(close);
                if (s < 0) return false;
                if (s == 0) return true;
            }
        }
        if (c == '"') {
            std::string s;
            return string(s);
        }
        ws();
        int k;
        while ((k = peek()) != EOF && k != ',' && k != '}' && k != ']' && !std::isspace(k)) get();
        return true;
    }

private:
    FILE* f_;
    char buf_[1 << 16];
    size_t pos_ = 0, len_ = 0;
    uint64_t read_ = 0;  // bytes read from the file into buf_ so far
};

// iterate the members of a JSON object: fn(key) parses the value
template <typename Fn>
bool json_object(Reader& r, Fn fn)
{
    if (!r.expect('{')) return false;
    r.ws();
    if (r.peek() == '}') {
        r.get();
        return true;
    }
    for (;;) {
        std::string key;
        if (!r.string(key) || !r.expect(':')) return false;
        if (!fn(key)) return false;
        const int s = r.sep('}');
        if (s < 0) return false;
        if (s == 0) return true;
    }
}

// iterate the elements of a JSON array: fn(index) parses the element
template <typename Fn>
bool json_array(Reader& r, Fn fn)
{
    if (!r.expect('[')) return false;
    r.ws();
    if (r.peek() == ']') {
        r.get();
        return true;
    }
    for (size_t i = 0;; ++i) {
        if (!fn(i)) return false;
        const int s = r.sep(']');
        if (s < 0) return false;
        if (s == 0) return true;
    }
}

bool read_state_json(Reader& r, HostState& s, bool& have_lattice, bool& have_beta, bool& have_links)
{
    return json_object(r, [&](const std::string& key) {
        if (key == "lattice") {
            have_lattice = true;
            return json_object(r, [&](const std::string& k) {
                if (k == "size") return r.number(s.size);
                if (k == "dim") return r.integer(s.points);
                return r.skip();
            });
        }
        if (key == "beta") {
            have_beta = true;
            return r.number(s.beta);
        }
        if (key == "link_matrix") {
            have_links = true;
            return json_object(r, [&](const std::string& k) {
                if (k != "data") return r.skip();
                s.links.clear();
                return json_array(r, [&](size_t) {
                    size_t n = 0;
                    const bool ok = json_array(r, [&](size_t) {
                        double re, im;
                        if (!r.expect('[') || !r.number(re) || r.sep(']') != 1 || !r.number(im) ||
                            r.sep(']') != 0)
                            return false;
                        s.links.push_back(re);
                        s.links.push_back(im);
                        ++n;
                        return n <= 9;
                    });
                    return ok && n == 9;
                });
            });
        }
        return r.skip();
    });
}

bool read_efield_json(Reader& r, HostState& s, int dim)
{
    return json_object(r, [&](const std::string& key) {
        if (key != "data") return r.skip();
        s.efield.clear();
        return json_array(r, [&](size_t) {
            int mu = 0;
            const bool ok = json_array(r, [&](size_t) {
                bool got = false;
                const bool o = json_object(r, [&](const std::string& k) {
                    if (k != "data") return r.skip();
                    got = true;
                    int a = 0;
                    const bool oa = json_array(r, [&](size_t) {
                        double v;
                        if (!r.number(v)) return false;
                        s.efield.push_back(v);
                        return ++a <= 8;
                    });
                    return oa && a == 8;
                });
                return o && got && ++mu <= dim;
            });
            return ok && mu == dim;
        });
    });
}

uint64_t ipow(uint64_t b, int e)
{
    uint64_t r = 1;
    for (int i = 0; i < e; ++i) {
        if (b && r > UINT64_MAX / b) return 0;
        r *= b;
    }
    return r;
}

// parse a file into s; with_e selects LatticeStateEFSyncDefault
int read_file(const char* path, int format, int with_e, int dim, HostState& s)
{
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return io_fail(std::string("cannot open for reading: ") + std::strerror(errno), path);
    // the file size bounds every length prefix before anything is allocated
    uint64_t fsize = 0;
    {
        struct stat sb;
        if (fstat(fileno(fh.f), &sb) != 0) return io_fail(std::string("cannot stat: ") + std::strerror(errno), path);
        fsize = (uint64_t)sb.st_size;
    }
    Reader r(fh.f);
    s.dim = dim;
    if (format == LQCD_FORMAT_BINCODE) {
        uint64_t n = 0;
        if (with_e) {
            if (!r.u64(n)) return io_fail("truncated E field length", path);
            // n sites of dim Su3Adjoint (8 doubles each)
            if (n > r.remaining(fsize) / ((uint64_t)dim * 64)) return io_fail("E field length exceeds the file", path);
            s.efield.resize(n * (size_t)dim * 8);
            if (!r.bytes(s.efield.data(), s.efield.size() * sizeof(double)))
                return io_fail("truncated E field", path);
            if (!r.u64(s.t)) return io_fail("truncated t", path);
        }
        if (!r.f64(s.size) || !r.u64(s.points) || !r.f64(s.beta) || !r.u64(n))
            return io_fail("truncated state header", path);
        // n matrices of 18 doubles
        if (n > r.remaining(fsize) / 144) return io_fail("link count exceeds the file", path);
        s.links.resize(n * 18);
        if (!r.bytes(s.links.data(), s.links.size() * sizeof(double))) return io_fail("truncated links", path);
        if (!r.at_end()) return io_fail("trailing bytes after the state", path);
    } else if (format == LQCD_FORMAT_JSON) {
        bool hl = false, hb = false, hk = false, he = !with_e, ht = !with_e, hs = !with_e;
        bool ok;
        if (with_e) {
            ok = json_object(r, [&](const std::string& key) {
                if (key == "e_field") {
                    he = true;
                    return read_efield_json(r, s, dim);
                }
                if (key == "t") {
                    ht = true;
                    return r.integer(s.t);
                }
                if (key == "lattice_state") {
                    hs = true;
                    return read_state_json(r, s, hl, hb, hk);
                }
                return r.skip();
            });
        } else {
            ok = read_state_json(r, s, hl, hb, hk);
        }
        if (!ok) return io_fail("malformed JSON for the expected serde layout", path);
        if (!(hl && hb && hk && he && ht && hs)) return io_fail("missing field in the JSON state", path);
        r.ws();
        if (!r.at_end()) return io_fail("trailing characters after the JSON state", path);
    } else {
        return fail(LQCD_ERR_INVALID_ARGUMENT, "unknown state format");
    }
    const uint64_t nsite = ipow(s.points, dim);
    if (s.points < 2 || nsite == 0)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "state lattice has fewer than 2 points per direction");
    if (s.links.size() != nsite * (uint64_t)dim * 18)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE,
                    "size of lattice and data are incompatible: " + std::to_string(s.links.size() / 18) +
                        " links for " + std::to_string(s.points) + "^" + std::to_string(dim) + " sites");
    if (with_e && s.efield.size() != nsite * (uint64_t)dim * 8)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "E field length does not match the lattice");
    return LQCD_OK;
}

int write_file(lqcd_lattice* L, double beta, int with_e, uint64_t t, const char* path, int format)
{
    CHECK_ARG(L && path, "null argument");
    CHECK_ARG(format == LQCD_FORMAT_BINCODE || format == LQCD_FORMAT_JSON, "unknown state format");
    for (int d = 1; d < L->g.dim; ++d)
        if (L->g.extent[d] != L->g.extent[0])
            return fail(LQCD_ERR_UNSUPPORTED,
                        "the reference's LatticeCyclic is hypercubic: extents must all be equal");
    HostState s;
    s.dim = L->g.dim;
    s.size = L->size;
    s.points = (uint64_t)L->g.extent[0];
    s.beta = beta;
    s.t = t;
    s.links.resize((size_t)L->nlink * 18);
    if (int rc = lqcd_links_download(L, s.links.data(), L->nlink)) return rc;
    if (with_e) {
        s.efield.resize((size_t)L->nlink * 8);
        if (int rc = lqcd_efield_download(L, s.efield.data(), L->nlink)) return rc;
    }
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return io_fail(std::string("cannot open for writing: ") + std::strerror(errno), path);
    {
        Writer w(fh.f);
        if (format == LQCD_FORMAT_BINCODE) {
            if (with_e) {
                write_efield_bincode(w, s);
                w.u64(t);
            }
            write_state_bincode(w, s);
        } else if (with_e) {
            w.text("{\"e_field\":");
            write_efield_json(w, s);
            w.text(",\"t\":");
            w.json_u64(t);
            w.text(",\"lattice_state\":");
            write_state_json(w, s);
            w.text("}");
        } else {
            write_state_json(w, s);
        }
        if (!w.ok()) return io_fail("write failed", path);
    }
    if (std::fclose(fh.f) != 0) {
        fh.f = nullptr;
        return io_fail("close failed", path);
    }
    fh.f = nullptr;
    return LQCD_OK;
}

} // namespace

extern "C" {

int lqcd_state_write(lqcd_lattice* L, double beta, const char* path, int format)
{
    return guarded([&] { return write_file(L, beta, 0, 0, path, format); });
}

int lqcd_state_ef_write(lqcd_lattice* L, double beta, uint64_t t, const char* path, int format)
{
    return guarded([&] { return write_file(L, beta, 1, t, path, format); });
}

int lqcd_state_read_header(const char* path, int format, int with_efield, int dim, double* size,
                           size_t* points, double* beta, uint64_t* t)
{
    CHECK_ARG(path && size && points && beta, "null argument");
    CHECK_ARG(dim >= 1 && dim <= 8, "dim must be in 1..8");
    HostState s;
    if (int rc = guarded([&] { return read_file(path, format, with_efield, dim, s); })) return rc;
    *size = s.size;
    *points = (size_t)s.points;
    *beta = s.beta;
    if (t) *t = s.t;
    return LQCD_OK;
}

int lqcd_state_read_host(const char* path, int format, int with_efield, int dim, double* links, size_t n_links,
                         double* efield)
{
    CHECK_ARG(path && links && (efield || !with_efield), "null argument");
    CHECK_ARG(dim >= 1 && dim <= 8, "dim must be in 1..8");
    HostState s;
    if (int rc = guarded([&] { return read_file(path, format, with_efield, dim, s); })) return rc;
    if (s.links.size() != n_links * 18)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "expected " + std::to_string(s.links.size() / 18) + " links, got " +
                                                    std::to_string(n_links));
    std::memcpy(links, s.links.data(), s.links.size() * sizeof(double));
    if (with_efield) std::memcpy(efield, s.efield.data(), s.efield.size() * sizeof(double));
    return LQCD_OK;
}

int lqcd_state_read(lqcd_lattice* L, const char* path, int format, int with_efield, double* beta,
                    uint64_t* t)
{
    CHECK_ARG(L && path, "null argument");
    HostState s;
    if (int rc = guarded([&] { return read_file(path, format, with_efield, L->g.dim, s); })) return rc;
    for (int d = 0; d < L->g.dim; ++d)
        if ((uint64_t)L->g.extent[d] != s.points)
            return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "the state's lattice does not match the handle's extents");
    if (s.size != L->size)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "the state's lattice spacing does not match the handle's");
    if (int rc = lqcd_links_upload(L, s.links.data(), L->nlink)) return rc;
    if (with_efield)
        if (int rc = lqcd_efield_upload(L, s.efield.data(), L->nlink)) return rc;
    if (beta) *beta = s.beta;
    if (t) *t = s.t;
    return LQCD_OK;
}

} // extern "C"
