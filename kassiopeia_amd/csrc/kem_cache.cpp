// kem_cache.cpp -- the charge-density solution cache of KEMField's charge-density solvers
// (host code, no device work).
//
// A solved container is stored in a directory of .kbd files and found again by hashes of its
// geometry, as KChargeDensitySolver::FindSolution / SaveSolution do
// (Interface/ChargeDensitySolvers/Electric/src/KChargeDensitySolver.cc:51-280):
//
// * hashes: MD5 (RFC 1321) over the container streamed as KMD5HashGenerator streams it
//   (IO/HashGenerator/include/KMD5HashGenerator.hh): integers as raw little-endian bytes, each
//   double rounded half up at its highest masked bit and the masked bits cleared, doubles of
//   magnitude below the threshold left out, and the omitted surface policies (the charge
//   densities; the Dirichlet boundary values) left out;
// * files: the KEMFile layout (IO/FileManipulation/include/KEMFile.hh) written by the raw binary
//   streamer (IO/Streamers/include/KBinaryDataStreamer.hh): per stored object a key (object name,
//   class name, object hash, labels, object location and size) followed by the object;
// * lookup: KEMFileInterface (IO/FileManipulation/include/KEMFileInterface.hh:23-81,
//   src/KEMFileInterface.cc): every .kbd file of the directory in name order, objects counted
//   and indexed by label sets across the files, containers found by hash;
// * a stored solution is reused when its residual threshold is at most the caller's tolerance;
//   otherwise stored solutions of the same shape with other boundary values are superposed when
//   their boundary vectors span the caller's (LinearAlgebra/Solvers/include/KSuperpositionSolver.hh).
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <set>
#include <string>
#include <vector>

#include "../../include/kemhip.h"

namespace {

thread_local std::string g_error;

int cache_fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

// ------------------------------------------------------------------ MD5 (RFC 1321)
class Md5 {
  public:
    Md5() { reset(); }
    void reset()
    {
        h_[0] = 0x67452301u;
        h_[1] = 0xefcdab89u;
        h_[2] = 0x98badcfeu;
        h_[3] = 0x10325476u;
        len_ = 0;
        fill_ = 0;
    }
    void update(const void* data, size_t n)
    {
        const uint8_t* p = (const uint8_t*)data;
        len_ += n;
        while (n > 0) {
            const size_t k = std::min(n, (size_t)64 - fill_);
            std::memcpy(buf_ + fill_, p, k);
            fill_ += k;
            p += k;
            n -= k;
            if (fill_ == 64) {
                block(buf_);
                fill_ = 0;
            }
        }
    }
    // finalises (padding: 0x80, zeros to 56 mod 64, the bit length little-endian) and returns
    // the 16 digest bytes, A B C D little-endian, as 32 lowercase hex digits
    std::string hex()
    {
        const uint64_t bits = len_ * 8;
        const uint8_t one = 0x80, zero = 0;
        update(&one, 1);
        while (fill_ != 56) update(&zero, 1);
        uint8_t lb[8];
        for (int i = 0; i < 8; i++) lb[i] = (uint8_t)(bits >> (8 * i));
        update(lb, 8);
        static const char* digits = "0123456789abcdef";
        std::string s;
        for (int w = 0; w < 4; w++)
            for (int b = 0; b < 4; b++) {
                const uint8_t v = (uint8_t)(h_[w] >> (8 * b));
                s += digits[v >> 4];
                s += digits[v & 15];
            }
        return s;
    }

  private:
    static uint32_t rotl(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }
    // T[i] = floor(2^32 |sin(i + 1)|) (RFC 1321, section 3.4)
    static const uint32_t* table()
    {
        static uint32_t t[64];
        static bool ready = false;
        if (!ready) {
            for (int i = 0; i < 64; i++) t[i] = (uint32_t)std::floor(std::fabs(std::sin((double)(i + 1))) * 4294967296.0);
            ready = true;
        }
        return t;
    }
    void block(const uint8_t* b)
    {
        static const int shift[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
        const uint32_t* T = table();
        uint32_t X[16];
        for (int i = 0; i < 16; i++)
            X[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
                   ((uint32_t)b[4 * i + 3] << 24);
        uint32_t a = h_[0], bb = h_[1], c = h_[2], d = h_[3];
        for (int i = 0; i < 64; i++) {
            const int round = i / 16;
            uint32_t f;
            int k;
            switch (round) {
                case 0: f = (bb & c) | (~bb & d); k = i; break;
                case 1: f = (bb & d) | (c & ~d); k = (5 * i + 1) & 15; break;
                case 2: f = bb ^ c ^ d; k = (3 * i + 5) & 15; break;
                default: f = c ^ (bb | ~d); k = (7 * i) & 15; break;
            }
            const uint32_t t = d;
            d = c;
            c = bb;
            bb = bb + rotl(a + f + X[k] + T[i], shift[round][i & 3]);
            a = t;
        }
        h_[0] += a;
        h_[1] += bb;
        h_[2] += c;
        h_[3] += d;
    }
    uint32_t h_[4];
    uint64_t len_;
    uint8_t buf_[64];
    size_t fill_;
};

// ------------------------------------------------------------------ streams
// A hash stream with KMD5HashGenerator's rules.  masked_bits = 8 * bytes + bits: the rounding
// bit is the highest masked one (bit bits - 1 of byte `bytes`, or the top bit of byte bytes - 1
// when bits = 0), and bytes 0 .. 6 are cleared below it (byte 7, the sign and exponent byte,
// never).
struct HashStream {
    Md5 md5;
    int masked_bits = 1;
    double threshold = 1e-12;
    void raw(const void* p, size_t n) { md5.update(p, n); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u16(uint16_t v) { raw(&v, 2); }
    void f64(double x)
    {
        if (std::fabs(x) < threshold) return;
        double y = x;
        unsigned char* yb = reinterpret_cast<unsigned char*>(&y);
        const int bits = masked_bits % 8, bytes = masked_bits / 8;
        if (bits || bytes) {
            const int index = bits ? bytes : bytes - 1;
            const unsigned char sig = bits ? (unsigned char)(1u << (bits - 1)) : (unsigned char)0x80;
            if (index < 8 && (yb[index] & sig)) {
                double round_off = y;
                reinterpret_cast<unsigned char*>(&round_off)[index] ^= sig;
                y += (y - round_off);
            }
        }
        for (int i = 0; i != 7; i++) {
            if (i == bytes) {
                yb[i] &= (unsigned char)(0xffu << bits);
                break;
            }
            yb[i] = 0;
        }
        md5.update(&y, 8);
    }
};

// raw little-endian binary (KBinaryDataStreamer): fundamental types as their bytes, strings as
// a 32-bit length and the characters
struct ByteWriter {
    std::string out;
    void raw(const void* p, size_t n) { out.append((const char*)p, n); }
    void u32(uint32_t v) { raw(&v, 4); }
    void u16(uint16_t v) { raw(&v, 2); }
    void u64(uint64_t v) { raw(&v, 8); }
    void f64(double v) { raw(&v, 8); }
    void str(const std::string& s)
    {
        u32((uint32_t)s.size());
        raw(s.data(), s.size());
    }
};

struct ByteReader {
    const std::string& in;
    size_t pos = 0;
    bool ok = true;
    explicit ByteReader(const std::string& s, size_t p = 0) : in(s), pos(p) {}
    bool take(void* p, size_t n)
    {
        if (!ok || pos + n > in.size()) return ok = false;
        std::memcpy(p, in.data() + pos, n);
        pos += n;
        return true;
    }
    uint32_t u32() { uint32_t v = 0; take(&v, 4); return v; }
    uint16_t u16() { uint16_t v = 0; take(&v, 2); return v; }
    uint64_t u64() { uint64_t v = 0; take(&v, 8); return v; }
    double f64() { double v = 0; take(&v, 8); return v; }
    std::string str()
    {
        const uint32_t n = u32();
        if (!ok || pos + n > in.size()) {
            ok = false;
            return std::string();
        }
        std::string s = in.substr(pos, n);
        pos += n;
        return s;
    }
};

// ------------------------------------------------------------------ surfaces
// doubles of each shape's record as its streamer writes them (KTriangle / KRectangle: A, B,
// P0, N1, N2; KLineSegment: P0, P1, diameter; KConicSection: P0, P1; KRing: P)
int shape_doubles(int shape)
{
    static const int nd[5] = {11, 11, 7, 6, 3};
    return (shape >= 0 && shape < 5) ? nd[shape] : -1;
}

struct Surfaces {
    int64_t n = 0;
    std::vector<int32_t> shape, bc;
    std::vector<double> data;  // n x 11
    std::vector<double> value, flux, sigma;
};

int check_surfaces(int64_t n, const int32_t* shape, const double* data, const int32_t* bc)
{
    if (n <= 0 || !shape || !data) return cache_fail(KEM_E_ARG, "cache: no surfaces");
    for (int64_t i = 0; i < n; i++) {
        if (shape_doubles(shape[i]) < 0) return cache_fail(KEM_E_ARG, "cache: unknown shape type");
        if (bc && bc[i] != KEM_BC_DIRICHLET && bc[i] != KEM_BC_NEUMANN)
            return cache_fail(KEM_E_ARG, "cache: unknown boundary type");
    }
    return KEM_OK;
}

Surfaces make_surfaces(int64_t n, const int32_t* shape, const double* data, const int32_t* bc, const double* value,
                       const double* flux, const double* sigma)
{
    Surfaces s;
    s.n = n;
    s.shape.assign(shape, shape + n);
    s.data.assign(data, data + 11 * n);
    s.bc.assign((size_t)n, KEM_BC_DIRICHLET);
    if (bc) s.bc.assign(bc, bc + n);
    s.value.assign((size_t)n, 0.0);
    if (value) s.value.assign(value, value + n);
    s.flux.assign((size_t)n, 0.0);
    if (flux) s.flux.assign(flux, flux + n);
    s.sigma.assign((size_t)n, 0.0);
    if (sigma) s.sigma.assign(sigma, sigma + n);
    return s;
}

constexpr int kOmitBasis = KEM_HASH_OMIT_BASIS, kOmitDirichlet = KEM_HASH_OMIT_DIRICHLET;

// KSurfaceContainer's streamer (Surfaces/include/KSurfaceContainer.hh:259-264, 495-539): the
// number of surface arrays, then per array its size, the KSurfaceID of its policies (basis,
// boundary, shape: unsigned shorts, indices into KBasisTypes / KBoundaryTypes / KShapeTypes)
// and each surface as shape, boundary, basis (Surfaces/include/KSurface.hh:164-173).  The
// arrays are the runs of equal (boundary, shape) types in container order: the container keeps
// one array per policy combination, in first-seen order, and its global order walks them.
template <typename S>
void stream_surfaces(S& s, const Surfaces& c, int omit, bool hash)
{
    std::vector<std::pair<int64_t, int64_t>> runs;
    for (int64_t i = 0; i < c.n;) {
        int64_t j = i + 1;
        while (j < c.n && c.shape[j] == c.shape[i] && c.bc[j] == c.bc[i]) j++;
        runs.push_back({i, j});
        i = j;
    }
    s.u32((uint32_t)runs.size());
    for (const auto& r : runs) {
        s.u32((uint32_t)(r.second - r.first));
        s.u16(0);  // KElectrostaticBasis
        s.u16((uint16_t)c.bc[r.first]);
        s.u16((uint16_t)c.shape[r.first]);
        const int nd = shape_doubles(c.shape[r.first]);
        const bool dirichlet = c.bc[r.first] == KEM_BC_DIRICHLET;
        for (int64_t i = r.first; i < r.second; i++) {
            for (int k = 0; k < nd; k++) s.f64(c.data[11 * i + k]);
            if (hash && dirichlet && (omit & kOmitDirichlet)) {
                // KBoundaryType<KElectrostaticBasis, KDirichletBoundary> omitted
            } else {
                s.f64(dirichlet ? c.value[i] : c.flux[i]);
            }
            if (!(hash && (omit & kOmitBasis))) s.f64(c.sigma[i]);
        }
    }
}

std::string hash_surfaces(const Surfaces& c, int omit, int masked_bits, double threshold)
{
    HashStream h;
    h.masked_bits = masked_bits;
    h.threshold = threshold;
    stream_surfaces(h, c, omit, true);
    return h.md5.hex();
}

// the container's binary form back into surfaces
bool read_surfaces(ByteReader& r, Surfaces& c)
{
    const uint32_t narrays = r.u32();
    c = Surfaces();
    for (uint32_t a = 0; a < narrays && r.ok; a++) {
        const uint32_t size = r.u32();
        const uint16_t basis = r.u16(), boundary = r.u16(), shape = r.u16();
        if (!r.ok || basis != 0 || boundary > 1 || shape_doubles(shape) < 0) return false;
        const int nd = shape_doubles(shape);
        for (uint32_t i = 0; i < size && r.ok; i++) {
            double rec[11] = {0};
            for (int k = 0; k < nd; k++) rec[k] = r.f64();
            const double bval = r.f64();
            const double sg = r.f64();
            c.shape.push_back(shape);
            c.bc.push_back(boundary);
            c.data.insert(c.data.end(), rec, rec + 11);
            c.value.push_back(boundary == KEM_BC_DIRICHLET ? bval : 0.0);
            c.flux.push_back(boundary == KEM_BC_NEUMANN ? bval : 0.0);
            c.sigma.push_back(sg);
        }
    }
    c.n = (int64_t)c.shape.size();
    return r.ok;
}

// ------------------------------------------------------------------ KEMFile
struct Key {
    std::string name, cls, hash;
    std::vector<std::string> labels;
    uint64_t location = 0, size = 0;
    size_t at = 0;  // offset of the key in the file
};

void write_key(ByteWriter& w, const Key& k)
{
    w.str(k.name);
    w.str(k.cls);
    w.str(k.hash);
    w.u32((uint32_t)k.labels.size());
    for (const auto& l : k.labels) w.str(l);
    w.u64(k.location);
    w.u64(k.size);
}

bool read_file(const std::string& path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

// the keys of a file in order; a key whose object does not lie past it ends the scan (the
// reference reports such a file as corrupted and stops reading it)
std::vector<Key> read_keys(const std::string& bytes)
{
    std::vector<Key> keys;
    size_t at = 0;
    while (at < bytes.size()) {
        ByteReader r(bytes, at);
        Key k;
        k.at = at;
        k.name = r.str();
        k.cls = r.str();
        k.hash = r.str();
        const uint32_t nl = r.u32();
        for (uint32_t i = 0; i < nl && r.ok; i++) k.labels.push_back(r.str());
        k.location = r.u64();
        k.size = r.u64();
        if (!r.ok) break;
        const uint64_t next = k.location + k.size;
        if (next <= at || next > bytes.size()) break;
        keys.push_back(k);
        at = (size_t)next;
    }
    return keys;
}

bool has_labels(const Key& k, const std::vector<std::string>& labels)
{
    for (const auto& l : labels)
        if (std::find(k.labels.begin(), k.labels.end(), l) == k.labels.end()) return false;
    return true;
}

// KEMFile::Write: skipped when an object of that name exists; otherwise the key (location and
// size completed after the object) and the object are appended
int file_write(const std::string& path, const std::string& name, const std::string& cls, const std::string& hash,
               const std::vector<std::string>& labels, const std::string& object, bool* written)
{
    *written = false;
    std::string bytes;
    if (read_file(path, bytes))
        for (const auto& k : read_keys(bytes))
            if (k.name == name) return KEM_OK;
    Key k;
    k.name = name;
    k.cls = cls;
    k.hash = hash;
    k.labels = labels;
    ByteWriter probe;
    write_key(probe, k);
    k.location = bytes.size() + probe.out.size();
    k.size = object.size();
    ByteWriter w;
    write_key(w, k);
    w.raw(object.data(), object.size());
    std::ofstream f(path, std::ios::binary | std::ios::app);
    if (!f) return cache_fail(KEM_E_ARG, "cache: cannot write " + path);
    f.write(w.out.data(), (std::streamsize)w.out.size());
    if (!f) return cache_fail(KEM_E_ARG, "cache: write failed on " + path);
    *written = true;
    return KEM_OK;
}

// ------------------------------------------------------------------ KEMFileInterface
struct Directory {
    std::string dir;
    std::set<std::string> files() const  // the directory's .kbd files, in name order
    {
        std::set<std::string> out;
        DIR* d = opendir(dir.c_str());
        if (!d) return out;
        while (dirent* e = readdir(d)) {
            const std::string s(e->d_name);
            const size_t dot = s.find_last_of('.');
            if (dot != std::string::npos && s.substr(dot) == ".kbd") out.insert(dir + "/" + s);
        }
        closedir(d);
        return out;
    }
    struct Hit {
        std::string file;
        std::string bytes;
        Key key;
    };
    // objects whose keys carry every label, across the files in order
    std::vector<Hit> labeled(const std::vector<std::string>& labels) const
    {
        std::vector<Hit> out;
        for (const auto& f : files()) {
            std::string bytes;
            if (!read_file(f, bytes)) continue;
            for (const auto& k : read_keys(bytes))
                if (has_labels(k, labels)) out.push_back({f, bytes, k});
        }
        return out;
    }
    // the first object, in file order, whose name or object hash is `hash` (KEMFile::HasElement
    // then ReadHashed, which reads the first key whose object hash matches)
    bool by_hash(const std::string& hash, const std::string& cls, Hit& hit) const
    {
        for (const auto& f : files()) {
            std::string bytes;
            if (!read_file(f, bytes)) continue;
            const std::vector<Key> keys = read_keys(bytes);
            bool has = false;
            for (const auto& k : keys) has = has || k.name == hash || k.hash == hash;
            if (!has) continue;
            for (const auto& k : keys)
                if (k.hash == hash) {
                    if (k.cls != cls) return false;
                    hit = {f, bytes, k};
                    return true;
                }
            return false;
        }
        return false;
    }
};

const char* kThresholdClass = "ResidualThreshold";
const char* kContainerClass = "SurfaceContainer";

struct Threshold {
    double value = std::numeric_limits<double>::max();
    std::string geometry_hash;
};

bool read_threshold(const Directory::Hit& h, Threshold& t)
{
    ByteReader r(h.bytes, (size_t)h.key.location);
    t.value = r.f64();
    t.geometry_hash = r.str();
    return r.ok;
}

bool read_container(const Directory::Hit& h, Surfaces& c)
{
    ByteReader r(h.bytes, (size_t)h.key.location);
    return read_surfaces(r, c);
}

// boundary vector b (KBoundaryIntegralVector): the Dirichlet value, 0 on Neumann rows
std::vector<double> boundary_vector(const Surfaces& c)
{
    std::vector<double> b((size_t)c.n);
    for (int64_t i = 0; i < c.n; i++) b[i] = c.bc[i] == KEM_BC_DIRICHLET ? c.value[i] : 0.0;
    return b;
}

// Least-squares weights w of A w = b (A: m x k, column-major, m >= k after padding) by a
// one-sided Jacobi SVD and its pseudo-inverse (singular values below 1e-14 of the largest
// dropped)
void svd_lstsq(std::vector<double> A, int m, int k, const std::vector<double>& b, std::vector<double>& w)
{
    // columns of A are rotated until orthogonal; V accumulates the rotations
    std::vector<double> V((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++) V[(size_t)i * k + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < k; p++)
            for (int q = p + 1; q < k; q++) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < m; i++) {
                    const double ap = A[(size_t)p * m + i], aq = A[(size_t)q * m + i];
                    alpha += ap * ap;
                    beta += aq * aq;
                    gamma += ap * aq;
                }
                if (gamma == 0.0) continue;
                off = std::max(off, std::fabs(gamma) / std::sqrt(alpha * beta));
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
                for (int i = 0; i < m; i++) {
                    const double ap = A[(size_t)p * m + i], aq = A[(size_t)q * m + i];
                    A[(size_t)p * m + i] = cs * ap - sn * aq;
                    A[(size_t)q * m + i] = sn * ap + cs * aq;
                }
                for (int i = 0; i < k; i++) {
                    const double vp = V[(size_t)p * k + i], vq = V[(size_t)q * k + i];
                    V[(size_t)p * k + i] = cs * vp - sn * vq;
                    V[(size_t)q * k + i] = sn * vp + cs * vq;
                }
            }
        if (off < 1e-15) break;
    }
    std::vector<double> s(k);
    double smax = 0;
    for (int j = 0; j < k; j++) {
        double nn = 0;
        for (int i = 0; i < m; i++) nn += A[(size_t)j * m + i] * A[(size_t)j * m + i];
        s[j] = std::sqrt(nn);
        smax = std::max(smax, s[j]);
    }
    w.assign(k, 0.0);
    for (int j = 0; j < k; j++) {
        if (s[j] <= smax * 1e-14 || s[j] == 0.0) continue;
        double ub = 0;  // u_j . b with u_j = A_j / s_j
        for (int i = 0; i < m; i++) ub += A[(size_t)j * m + i] * b[i];
        ub /= s[j] * s[j];
        for (int i = 0; i < k; i++) w[i] += V[(size_t)j * k + i] * ub;
    }
}

}  // namespace

extern "C" {

const char* kem_cache_error(void) { return g_error.c_str(); }

int kem_md5_hex(const void* data, int64_t bytes, char* hex33)
{
    if (!hex33 || bytes < 0 || (bytes > 0 && !data)) return cache_fail(KEM_E_ARG, "md5: bad arguments");
    Md5 m;
    m.update(data, (size_t)bytes);
    std::memcpy(hex33, m.hex().c_str(), 33);
    return KEM_OK;
}

int kem_hash_doubles(const double* x, int64_t n, int masked_bits, double threshold, char* hex33)
{
    if (!hex33 || n < 0 || (n > 0 && !x) || masked_bits < 0 || masked_bits > 63)
        return cache_fail(KEM_E_ARG, "hash: bad arguments");
    HashStream h;
    h.masked_bits = masked_bits;
    h.threshold = threshold;
    for (int64_t i = 0; i < n; i++) h.f64(x[i]);
    std::memcpy(hex33, h.md5.hex().c_str(), 33);
    return KEM_OK;
}

int kem_hash_surfaces(int64_t n, const int32_t* shape_type, const double* shape_data, const int32_t* bc_type,
                      const double* bc_value, const double* bc_flux, const double* sigma, int omit, int masked_bits,
                      double threshold, char* hex33)
{
    if (!hex33 || masked_bits < 0 || masked_bits > 63) return cache_fail(KEM_E_ARG, "hash: bad arguments");
    const int rc = check_surfaces(n, shape_type, shape_data, bc_type);
    if (rc) return rc;
    const Surfaces c = make_surfaces(n, shape_type, shape_data, bc_type, bc_value, bc_flux, sigma);
    std::memcpy(hex33, hash_surfaces(c, omit, masked_bits, threshold).c_str(), 33);
    return KEM_OK;
}

int kem_cache_save(const char* directory, const char* file, int64_t n, const int32_t* shape_type,
                   const double* shape_data, const int32_t* bc_type, const double* bc_value, const double* bc_flux,
                   const double* sigma, double residual_threshold, int masked_bits, double hash_threshold)
{
    if (!directory || !file || !sigma || masked_bits < 0 || masked_bits > 63)
        return cache_fail(KEM_E_ARG, "cache save: bad arguments");
    const int rc = check_surfaces(n, shape_type, shape_data, bc_type);
    if (rc) return rc;
    const Surfaces c = make_surfaces(n, shape_type, shape_data, bc_type, bc_value, bc_flux, sigma);
    mkdir(directory, S_IRWXU);  // KEMFileInterface::ActiveDirectory creates a missing directory
    const std::string path = std::string(directory) + "/" + file;
    const std::string shape_hash = hash_surfaces(c, kOmitBasis | kOmitDirichlet, masked_bits, hash_threshold);
    const std::string boundary_hash = hash_surfaces(c, kOmitBasis, masked_bits, hash_threshold);
    // KMD5HashGenerator's defaults (1 masked bit, threshold 1e-12): the solution's own hash,
    // which is also the object hash of the stored container's key
    const std::string solution_hash = hash_surfaces(c, 0, 1, 1e-12);
    ByteWriter thr;
    thr.f64(residual_threshold);
    thr.str(solution_hash);
    // the threshold's object hash covers its value and the geometry hash's characters (the
    // reference hashes the string object's bytes; nothing looks thresholds up by hash)
    HashStream th;
    th.f64(residual_threshold);
    th.raw(solution_hash.data(), solution_hash.size());
    bool written = false;
    int r = file_write(path, std::string(kThresholdClass) + "_" + solution_hash, kThresholdClass, th.md5.hex(),
                       {kThresholdClass, shape_hash, boundary_hash, solution_hash}, thr.out, &written);
    if (r) return r;
    ByteWriter obj;
    stream_surfaces(obj, c, 0, false);
    return file_write(path, std::string(kContainerClass) + "_" + solution_hash, kContainerClass, solution_hash,
                      {kContainerClass, solution_hash}, obj.out, &written);
}

int kem_cache_find(const char* directory, int64_t n, const int32_t* shape_type, const double* shape_data,
                   const int32_t* bc_type, const double* bc_value, const double* bc_flux, double tolerance,
                   int masked_bits, double hash_threshold, double* sigma, int* found, char* file_out,
                   int64_t file_cap)
{
    if (!directory || !sigma || !found || masked_bits < 0 || masked_bits > 63)
        return cache_fail(KEM_E_ARG, "cache find: bad arguments");
    *found = 0;
    if (file_out && file_cap > 0) file_out[0] = 0;
    const int rc = check_surfaces(n, shape_type, shape_data, bc_type);
    if (rc) return rc;
    const Surfaces c = make_surfaces(n, shape_type, shape_data, bc_type, bc_value, bc_flux, nullptr);
    const Directory dir{directory};
    const std::string shape_hash = hash_surfaces(c, kOmitBasis | kOmitDirichlet, masked_bits, hash_threshold);
    const std::string boundary_hash = hash_surfaces(c, kOmitBasis, masked_bits, hash_threshold);
    auto report = [&](const std::string& f) {
        if (file_out && file_cap > 0) {
            std::strncpy(file_out, f.c_str(), (size_t)file_cap - 1);
            file_out[file_cap - 1] = 0;
        }
    };

    // a stored solution of this shape and these boundary values: the one of smallest threshold
    const auto exact = dir.labeled({kThresholdClass, shape_hash, boundary_hash});
    if (!exact.empty()) {
        Threshold best;
        for (const auto& h : exact) {
            Threshold t;
            if (read_threshold(h, t) && t.value < best.value) best = t;
        }
        Directory::Hit hc;
        Surfaces stored;
        if (best.value <= tolerance && dir.by_hash(best.geometry_hash, kContainerClass, hc) &&
            read_container(hc, stored) && stored.n == n) {
            std::copy(stored.sigma.begin(), stored.sigma.end(), sigma);
            *found = 1;
            report(hc.file);
            return KEM_OK;
        }
    }

    // solutions of this shape with other boundary values, superposed (KSuperpositionSolver):
    // rows with equal values across every stored system and the caller's are merged, and the
    // reduced system is solved in the least-squares sense; accepted when the residual's largest
    // entry is below 1e-14 (KSVDSolver.hh:97)
    const auto same_shape = dir.labeled({kThresholdClass, shape_hash});
    std::vector<std::vector<double>> xs, bs;
    std::string last_file;
    for (const auto& h : same_shape) {
        Threshold t;
        if (!read_threshold(h, t) || !(t.value <= tolerance)) continue;
        Directory::Hit hc;
        Surfaces stored;
        if (!dir.by_hash(t.geometry_hash, kContainerClass, hc) || !read_container(hc, stored) || stored.n != n)
            continue;
        xs.push_back(stored.sigma);
        bs.push_back(boundary_vector(stored));
        last_file = hc.file;
    }
    if (xs.empty()) return KEM_OK;
    const std::vector<double> b = boundary_vector(c);
    const size_t ns = xs.size();
    std::vector<std::vector<double>> red(ns + 1);  // reduced rows per system (last: the caller's)
    const double tol = 1e-14;
    for (int64_t i = 0; i < n; i++) {
        size_t k = 0;
        for (;; k++) {
            if (k == red[0].size()) {
                for (size_t j = 0; j < ns; j++) red[j].push_back(bs[j][i]);
                red[ns].push_back(b[i]);
            }
            // row i joins reduced row k when every system, the caller's included, has its value
            // there (the reference's early exit on the caller's vector can never trigger:
            // KSuperpositionSolver.hh:123 compares a loop index with the loop bound)
            bool match = std::fabs(red[ns][k] - b[i]) <= tol;
            for (size_t j = 0; j < ns && match; j++) match = std::fabs(red[j][k] - bs[j][i]) <= tol;
            if (match) break;
        }
    }
    const int m0 = (int)red[0].size(), k = (int)ns;
    const int m = std::max(m0, k);
    std::vector<double> A((size_t)m * k, 0.0), rb((size_t)m, 0.0);
    for (int j = 0; j < k; j++)
        for (int i = 0; i < m0; i++) A[(size_t)j * m + i] = red[j][i];
    for (int i = 0; i < m0; i++) rb[i] = red[ns][i];
    std::vector<double> w;
    svd_lstsq(A, m, k, rb, w);
    double resid = 0.0;
    for (int i = 0; i < m0; i++) {
        double s = 0;
        for (int j = 0; j < k; j++) s += red[j][i] * w[j];
        resid = std::max(resid, std::fabs(s - rb[i]));
    }
    if (!(resid < tol)) return KEM_OK;
    for (int64_t i = 0; i < n; i++) {
        double s = 0;
        for (size_t j = 0; j < ns; j++) s += w[j] * xs[j][i];
        sigma[i] = s;
    }
    *found = 2;
    report(last_file);
    return KEM_OK;
}

}  // extern "C"
