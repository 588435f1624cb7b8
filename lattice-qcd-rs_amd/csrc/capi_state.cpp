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
// the lengths against dim^D.  Links and E are the host layouts of
// lqcd_links_download / lqcd_efield_download, which are the same orders.
//
// bincode files are fixed-width fields and f64 blocks whose lengths are u64
// prefixes: any bytes of the right length decode, so checking each prefix
// against the file size checks the whole file.  Their blocks stream between
// the file and HBM through a ring of pinned host chunks (the device packs or
// unpacks the AoS order), with the file I/O of one chunk overlapping the copy
// of the next, and nothing is written into the lattice before the whole file
// has been read.  JSON files are parsed on the host.
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

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

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
    int get()
    {
        if (pos_ == len_) {
            len_ = std::fread(buf_, 1, sizeof(buf_), f_);
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

// the lengths of a parsed state against its lattice: points^dim sites, dim
// links per site, and (with E) one Su3Adjoint per link
int check_shape(int dim, uint64_t points, uint64_t n_links, int with_e, uint64_t n_e)
{
    const uint64_t nsite = ipow(points, dim);
    if (points < 2 || nsite == 0)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "state lattice has fewer than 2 points per direction");
    if (n_links != nsite * (uint64_t)dim)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE,
                    "size of lattice and data are incompatible: " + std::to_string(n_links) + " links for " +
                        std::to_string(points) + "^" + std::to_string(dim) + " sites");
    if (with_e && n_e != nsite * (uint64_t)dim)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "E field length does not match the lattice");
    return LQCD_OK;
}

// parse a JSON file into s; with_e selects LatticeStateEFSyncDefault
int read_json_file(const char* path, int with_e, int dim, HostState& s)
{
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return io_fail(std::string("cannot open for reading: ") + std::strerror(errno), path);
    Reader r(fh.f);
    s.dim = dim;
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
    return check_shape(dim, s.points, s.links.size() / 18, with_e, s.efield.size() / 8);
}

// ------------------------------------------------------------------ bincode

// POSIX file descriptor closed on scope exit
struct Fd {
    int fd = -1;
    ~Fd()
    {
        if (fd >= 0) ::close(fd);
    }
};

bool pread_all(int fd, void* p, size_t n, uint64_t off)
{
    char* c = static_cast<char*>(p);
    while (n) {
        const ssize_t k = ::pread(fd, c, n, (off_t)off);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) return false;
        c += k;
        n -= (size_t)k;
        off += (uint64_t)k;
    }
    return true;
}

bool write_all(int fd, const void* p, size_t n)
{
    const char* c = static_cast<const char*>(p);
    while (n) {
        const ssize_t k = ::write(fd, c, n);
        if (k < 0 && errno == EINTR) continue;
        if (k <= 0) return false;
        c += k;
        n -= (size_t)k;
    }
    return true;
}

// little-endian fixed-width fields
void put_u64(unsigned char* b, uint64_t v)
{
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
}
void put_f64(unsigned char* b, double v)
{
    uint64_t u;
    std::memcpy(&u, &v, 8);
    put_u64(b, u);
}
uint64_t get_u64(const unsigned char* b)
{
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)b[i] << (8 * i);
    return v;
}
double get_f64(const unsigned char* b)
{
    const uint64_t u = get_u64(b);
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

// where the fields and blocks of a bincode state file lie
struct BincodeLayout {
    double size = 0.0, beta = 0.0;
    uint64_t points = 0, t = 0;
    uint64_t e_off = 0, e_sites = 0;       // with E: e_sites x dim Su3Adjoint (8 f64) at e_off
    uint64_t links_off = 0, n_links = 0;   // n_links matrices of 18 f64 at links_off
};

// opens a bincode file, reads its fixed-width fields, checks every length
// prefix against the file size and the lengths against the lattice; the
// blocks themselves are not read
int open_bincode(const char* path, int with_e, int dim, Fd& f, BincodeLayout& b)
{
    f.fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (f.fd < 0) return io_fail(std::string("cannot open for reading: ") + std::strerror(errno), path);
    struct stat sb;
    if (fstat(f.fd, &sb) != 0) return io_fail(std::string("cannot stat: ") + std::strerror(errno), path);
    const uint64_t fsize = (uint64_t)sb.st_size;
    uint64_t pos = 0;
    unsigned char h[32];
    // n bytes of fields at pos, or the error `what`
    auto field = [&](size_t n, const char* what) {
        if (fsize - pos < n) return io_fail(what, path);
        if (!pread_all(f.fd, h, n, pos)) return io_fail(std::string("read failed: ") + std::strerror(errno), path);
        pos += n;
        return (int)LQCD_OK;
    };
    if (with_e) {
        if (int rc = field(8, "truncated E field length")) return rc;
        b.e_sites = get_u64(h);
        if (b.e_sites > (fsize - pos) / ((uint64_t)dim * 64)) return io_fail("E field length exceeds the file", path);
        b.e_off = pos;
        pos += b.e_sites * (uint64_t)dim * 64;
        if (int rc = field(8, "truncated t")) return rc;
        b.t = get_u64(h);
    }
    if (int rc = field(32, "truncated state header")) return rc;
    b.size = get_f64(h);
    b.points = get_u64(h + 8);
    b.beta = get_f64(h + 16);
    b.n_links = get_u64(h + 24);
    if (b.n_links > (fsize - pos) / 144) return io_fail("link count exceeds the file", path);
    b.links_off = pos;
    if (fsize - pos != b.n_links * 144) return io_fail("trailing bytes after the state", path);
    return check_shape(dim, b.points, b.n_links, with_e, b.e_sites * (uint64_t)dim);
}

// Pinned host chunks between a file and HBM.  A slot's event marks the last
// copy through it; the slot is reused only after that event.
class PinnedRing {
public:
    static constexpr int kSlots = 4;
    static constexpr size_t kChunk = size_t(8) << 20;
    ~PinnedRing()
    {
        for (int k = 0; k < kSlots; ++k) {
            if (ev_[k]) {
                (void)hipEventSynchronize(ev_[k]);
                (void)hipEventDestroy(ev_[k]);
            }
            if (buf_[k]) (void)hipHostFree(buf_[k]);
        }
    }
    hipError_t init()
    {
        for (int k = 0; k < kSlots; ++k) {
            if (hipError_t e = hipHostMalloc(&buf_[k], kChunk)) return e;
            if (hipError_t e = hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming)) return e;
        }
        return hipSuccess;
    }
    char* buf(int k) { return static_cast<char*>(buf_[k]); }
    hipEvent_t ev(int k) { return ev_[k]; }

private:
    void* buf_[kSlots] = {};
    hipEvent_t ev_[kSlots] = {};
};

// n bytes of the file at off into device memory at dst (enqueued on s)
int file_to_device(int fd, uint64_t off, uint64_t n, char* dst, hipStream_t s, PinnedRing& R, const char* path)
{
    for (uint64_t c = 0; c * PinnedRing::kChunk < n; ++c) {
        const int k = (int)(c % PinnedRing::kSlots);
        const uint64_t o = c * PinnedRing::kChunk;
        const size_t len = (size_t)std::min<uint64_t>(PinnedRing::kChunk, n - o);
        HIP_TRY(hipEventSynchronize(R.ev(k)));
        if (!pread_all(fd, R.buf(k), len, off + o))
            return io_fail(std::string("read failed: ") + std::strerror(errno), path);
        HIP_TRY(hipMemcpyAsync(dst + o, R.buf(k), len, hipMemcpyHostToDevice, s));
        HIP_TRY(hipEventRecord(R.ev(k), s));
    }
    return LQCD_OK;
}

// n bytes of device memory at src appended to the file, kSlots chunks in flight
int device_to_file(int fd, const char* src, uint64_t n, hipStream_t s, PinnedRing& R, const char* path)
{
    const uint64_t nchunk = (n + PinnedRing::kChunk - 1) / PinnedRing::kChunk;
    auto chunk_len = [&](uint64_t c) { return (size_t)std::min<uint64_t>(PinnedRing::kChunk, n - c * PinnedRing::kChunk); };
    uint64_t issued = 0;
    for (uint64_t c = 0; c < nchunk; ++c) {
        // slot c % kSlots was last read by the write of chunk c - kSlots, done
        for (; issued < nchunk && issued < c + PinnedRing::kSlots; ++issued) {
            const int k = (int)(issued % PinnedRing::kSlots);
            HIP_TRY(hipMemcpyAsync(R.buf(k), src + issued * PinnedRing::kChunk, chunk_len(issued),
                                   hipMemcpyDeviceToHost, s));
            HIP_TRY(hipEventRecord(R.ev(k), s));
        }
        const int k = (int)(c % PinnedRing::kSlots);
        HIP_TRY(hipEventSynchronize(R.ev(k)));
        if (!write_all(fd, R.buf(k), chunk_len(c)))
            return io_fail(std::string("write failed: ") + std::strerror(errno), path);
    }
    return LQCD_OK;
}

int write_bincode(lqcd_lattice* L, double beta, int with_e, uint64_t t, const char* path)
{
    DeviceGuard dg(L->ctx->device);
    hipStream_t s = L->ctx->stream;
    const uint64_t ne = with_e ? (uint64_t)L->nlink * 64 : 0, nl = (uint64_t)L->nlink * 144;
    DevBuf stage;
    HIP_TRY(hipMalloc(&stage.p, ne + nl));
    char* d = static_cast<char*>(stage.p);
    if (with_e) HIP_TRY(launch_efield_to_aos(L, L->E, (double*)d));
    HIP_TRY(launch_links_to_aos(L, L->U[L->cur], (double*)(d + ne)));
    PinnedRing R;
    HIP_TRY(R.init());
    Fd f;
    f.fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0666);
    if (f.fd < 0) return io_fail(std::string("cannot open for writing: ") + std::strerror(errno), path);
    unsigned char h[32];
    auto fields = [&](size_t n) {
        return write_all(f.fd, h, n) ? (int)LQCD_OK : io_fail(std::string("write failed: ") + std::strerror(errno), path);
    };
    if (with_e) {
        put_u64(h, (uint64_t)L->g.nsite);
        if (int rc = fields(8)) return rc;
        if (int rc = device_to_file(f.fd, d, ne, s, R, path)) return rc;
        put_u64(h, t);
        if (int rc = fields(8)) return rc;
    }
    put_f64(h, L->size);
    put_u64(h + 8, (uint64_t)L->g.extent[0]);
    put_f64(h + 16, beta);
    put_u64(h + 24, (uint64_t)L->nlink);
    if (int rc = fields(32)) return rc;
    if (int rc = device_to_file(f.fd, d + ne, nl, s, R, path)) return rc;
    const int fd = f.fd;
    f.fd = -1;
    if (::close(fd) != 0) return io_fail(std::string("close failed: ") + std::strerror(errno), path);
    return LQCD_OK;
}

int read_bincode_into(lqcd_lattice* L, const char* path, int with_e, double* beta, uint64_t* t)
{
    Fd f;
    BincodeLayout b;
    if (int rc = open_bincode(path, with_e, L->g.dim, f, b)) return rc;
    for (int d = 0; d < L->g.dim; ++d)
        if ((uint64_t)L->g.extent[d] != b.points)
            return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "the state's lattice does not match the handle's extents");
    if (b.size != L->size)
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE, "the state's lattice spacing does not match the handle's");
    DeviceGuard dg(L->ctx->device);
    hipStream_t s = L->ctx->stream;
    const uint64_t ne = with_e ? (uint64_t)L->nlink * 64 : 0, nl = (uint64_t)L->nlink * 144;
    DevBuf stage;
    HIP_TRY(hipMalloc(&stage.p, ne + nl));
    char* d = static_cast<char*>(stage.p);
    {
        PinnedRing R;
        HIP_TRY(R.init());
        if (with_e)
            if (int rc = file_to_device(f.fd, b.e_off, ne, d, s, R, path)) return rc;
        if (int rc = file_to_device(f.fd, b.links_off, nl, d + ne, s, R, path)) return rc;
    }
    // the whole file is in HBM: only now is the lattice overwritten
    HIP_TRY(launch_links_from_aos(L, (const double*)(d + ne), L->U[L->cur]));
    if (with_e) HIP_TRY(launch_efield_from_aos(L, (const double*)d, L->E));
    HIP_TRY(hipStreamSynchronize(s));
    L->su3_known = -1;
    if (beta) *beta = b.beta;
    if (t) *t = b.t;
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
    if (format == LQCD_FORMAT_BINCODE) return write_bincode(L, beta, with_e, t, path);
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
        if (with_e) {
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
    CHECK_ARG(format == LQCD_FORMAT_BINCODE || format == LQCD_FORMAT_JSON, "unknown state format");
    double sz, b;
    uint64_t pts, tt;
    if (format == LQCD_FORMAT_BINCODE) {
        Fd f;
        BincodeLayout l;
        if (int rc = guarded([&] { return open_bincode(path, with_efield, dim, f, l); })) return rc;
        sz = l.size, pts = l.points, b = l.beta, tt = l.t;
    } else {
        HostState s;
        if (int rc = guarded([&] { return read_json_file(path, with_efield, dim, s); })) return rc;
        sz = s.size, pts = s.points, b = s.beta, tt = s.t;
    }
    *size = sz;
    *points = (size_t)pts;
    *beta = b;
    if (t) *t = tt;
    return LQCD_OK;
}

int lqcd_state_read_host(const char* path, int format, int with_efield, int dim, double* links, size_t n_links,
                         double* efield)
{
    CHECK_ARG(path && links && (efield || !with_efield), "null argument");
    CHECK_ARG(dim >= 1 && dim <= 8, "dim must be in 1..8");
    CHECK_ARG(format == LQCD_FORMAT_BINCODE || format == LQCD_FORMAT_JSON, "unknown state format");
    auto count_mismatch = [&](uint64_t have) {
        return fail(LQCD_ERR_INCOMPATIBLE_SIZE,
                    "expected " + std::to_string(have) + " links, got " + std::to_string(n_links));
    };
    if (format == LQCD_FORMAT_BINCODE) {
        Fd f;
        BincodeLayout l;
        if (int rc = guarded([&] { return open_bincode(path, with_efield, dim, f, l); })) return rc;
        if (l.n_links != n_links) return count_mismatch(l.n_links);
        if (!pread_all(f.fd, links, n_links * 144, l.links_off) ||
            (with_efield && !pread_all(f.fd, efield, n_links * 64, l.e_off)))
            return io_fail(std::string("read failed: ") + std::strerror(errno), path);
        return LQCD_OK;
    }
    HostState s;
    if (int rc = guarded([&] { return read_json_file(path, with_efield, dim, s); })) return rc;
    if (s.links.size() != n_links * 18) return count_mismatch(s.links.size() / 18);
    std::memcpy(links, s.links.data(), s.links.size() * sizeof(double));
    if (with_efield) std::memcpy(efield, s.efield.data(), s.efield.size() * sizeof(double));
    return LQCD_OK;
}

int lqcd_state_read(lqcd_lattice* L, const char* path, int format, int with_efield, double* beta,
                    uint64_t* t)
{
    CHECK_ARG(L && path, "null argument");
    CHECK_ARG(format == LQCD_FORMAT_BINCODE || format == LQCD_FORMAT_JSON, "unknown state format");
    if (format == LQCD_FORMAT_BINCODE)
        return guarded([&] { return read_bincode_into(L, path, with_efield, beta, t); });
    HostState s;
    if (int rc = guarded([&] { return read_json_file(path, with_efield, L->g.dim, s); })) return rc;
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
