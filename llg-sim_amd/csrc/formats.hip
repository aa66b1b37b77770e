// On-disk formats of llg-sim, written host-side by libllgmi.so:
//
//  * OOMMF OVF 2.0 rectangular meshes, text and binary4 (ovf.rs:54-262), with
//    the reference's header lines, order and number formatting (Rust's `{:.Ne}`
//    prints the exponent without a sign for positive values and without
//    leading zeros: 1.5e-9, 2.0e0);
//  * the demag kernel cache "LLGDMAG\0" version 2 (demag_fft_uniform.rs:
//    206-429): a 68-byte little-endian header and the four k-space kernels
//    (complex f64, px*py each, row-major) of Demag2D at the reference's FFT
//    sizes (px = 2 nx on an open axis, nx on a periodic one).
//
// No GPU code here.  The kernel cache needs FFTs of the reference's lengths,
// which need not be powers of two; they run on the host (mixed radix 2/3/4/5
// with a direct DFT for other prime factors), once per cache read or write.
#include "../../include/llgmi.h"
#include "common.hpp"
#include "formats.hpp"

#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace llgmi {

// Rust `{:.{prec}e}` of a finite f64 (or NaN / inf as Rust prints them).
std::string rust_exp(double v, int prec) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, v);
    std::string s(buf);
    const size_t e = s.find('e');
    if (e == std::string::npos) return s;
    std::string mant = s.substr(0, e), ex = s.substr(e + 1);
    bool neg = false;
    size_t i = 0;
    if (i < ex.size() && (ex[i] == '+' || ex[i] == '-')) neg = ex[i++] == '-';
    while (i + 1 < ex.size() && ex[i] == '0') ++i;
    return mant + "e" + (neg ? "-" : "") + ex.substr(i);
}

// create_dir_all of the parent directory of path
void ensure_parent_dir(const std::string &path) {
    const size_t slash = path.find_last_of('/');
    if (slash == std::string::npos || slash == 0) return;
    std::string dir = path.substr(0, slash);
    for (size_t p = 1; p <= dir.size(); ++p) {
        if (p == dir.size() || dir[p] == '/') {
            std::string part = dir.substr(0, p);
            if (::mkdir(part.c_str(), 0777) != 0 && errno != EEXIST)
                throw Error(LLGMI_ERR_INVALID, "cannot create directory " + part + ": " + std::strerror(errno));
        }
    }
}

namespace {

struct File {
    std::FILE *f = nullptr;
    std::string path;
    explicit File(const std::string &p, const char *mode) : path(p) {
        f = std::fopen(p.c_str(), mode);
        if (!f) throw Error(LLGMI_ERR_INVALID, "cannot open " + p + ": " + std::strerror(errno));
    }
    ~File() {
        if (f) std::fclose(f);
    }
    void put(const std::string &s) {
        if (std::fwrite(s.data(), 1, s.size(), f) != s.size())
            throw Error(LLGMI_ERR_INVALID, "write failed: " + path);
    }
    void line(const std::string &s) { put(s + "\n"); }
    void bytes(const void *p, size_t n) {
        if (std::fwrite(p, 1, n, f) != n) throw Error(LLGMI_ERR_INVALID, "write failed: " + path);
    }
    void read(void *p, size_t n) {
        if (std::fread(p, 1, n, f) != n) throw Error(LLGMI_ERR_INVALID, "short read: " + path);
    }
    void close() {
        if (f && std::fclose(f) != 0) {
            f = nullptr;
            throw Error(LLGMI_ERR_INVALID, "close failed: " + path);
        }
        f = nullptr;
    }
};

struct Meta {
    std::string title = "m";
    std::vector<std::string> desc;
    std::string labels[3] = {"m_x", "m_y", "m_z"};
    std::string units[3] = {"1", "1", "1"};
};

// OvfMeta::magnetization() (ovf.rs:26-33) with the caller's fields where given
Meta meta_from(const llgmi_ovf_meta *m) {
    Meta o;
    if (!m) return o;
    if (m->title) o.title = m->title;
    for (int32_t i = 0; i < m->n_desc; ++i)
        if (m->desc_lines && m->desc_lines[i]) o.desc.emplace_back(m->desc_lines[i]);
    for (int q = 0; q < 3; ++q) {
        if (m->valuelabels[q]) o.labels[q] = m->valuelabels[q];
        if (m->valueunits[q]) o.units[q] = m->valueunits[q];
    }
    return o;
}

}  // namespace

// write_ovf2_rectangular_text (ovf.rs:54-159)
void write_ovf2_text(const std::string &path, const llgmi_grid &g, const double *m,
                     const llgmi_ovf_meta *meta_c) {
    const Meta meta = meta_from(meta_c);
    ensure_parent_dir(path);
    const int64_t nx = g.nx, ny = g.ny, nz = 1;
    File f(path, "wb");
    f.line("# OOMMF OVF 2.0");
    f.line("# Segment count: 1");
    f.line("# Begin: Segment");
    f.line("# Begin: Header");
    f.line("# Title: " + meta.title);
    f.line("# meshtype: rectangular");
    f.line("# meshunit: m");
    f.line("# xmin: " + rust_exp(0.0, 16));
    f.line("# ymin: " + rust_exp(0.0, 16));
    f.line("# zmin: " + rust_exp(0.0, 16));
    f.line("# xmax: " + rust_exp((double)nx * g.dx, 16));
    f.line("# ymax: " + rust_exp((double)ny * g.dy, 16));
    f.line("# zmax: " + rust_exp((double)nz * g.dz, 16));
    f.line("# valuedim: 3");
    f.line("# valuelabels: " + meta.labels[0] + " " + meta.labels[1] + " " + meta.labels[2]);
    f.line("# valueunits: " + meta.units[0] + " " + meta.units[1] + " " + meta.units[2]);
    for (auto &d : meta.desc) f.line("# Desc: " + d);
    f.line("# xbase: " + rust_exp(0.5 * g.dx, 16));
    f.line("# ybase: " + rust_exp(0.5 * g.dy, 16));
    f.line("# zbase: " + rust_exp(0.5 * g.dz, 16));
    f.line("# xnodes: " + std::to_string(nx));
    f.line("# ynodes: " + std::to_string(ny));
    f.line("# znodes: " + std::to_string(nz));
    f.line("# xstepsize: " + rust_exp(g.dx, 16));
    f.line("# ystepsize: " + rust_exp(g.dy, 16));
    f.line("# zstepsize: " + rust_exp(g.dz, 16));
    f.line("# End: Header");
    f.line("# Begin: Data Text");
    std::string buf;
    buf.reserve(1 << 16);
    for (int64_t c = 0; c < nx * ny; ++c) {
        buf += rust_exp(m[3 * c], 10);
        buf += ' ';
        buf += rust_exp(m[3 * c + 1], 10);
        buf += ' ';
        buf += rust_exp(m[3 * c + 2], 10);
        buf += '\n';
        if (buf.size() > (1 << 16) - 128) {
            f.put(buf);
            buf.clear();
        }
    }
    f.put(buf);
    f.line("# End: Data Text");
    f.line("# End: Segment");
    f.close();
}

// write_ovf2_rectangular_binary4 (ovf.rs:161-262)
void write_ovf2_binary4(const std::string &path, const llgmi_grid &g, const double *m,
                        const llgmi_ovf_meta *meta_c) {
    const Meta meta = meta_from(meta_c);
    ensure_parent_dir(path);
    const int64_t nx = g.nx, ny = g.ny, nz = 1;
    File f(path, "wb");
    f.line("# OOMMF OVF 2.0");
    f.line("# Segment count: 1");
    f.line("# Begin: Segment");
    f.line("# Begin: Header");
    f.line("# Title: " + meta.title);
    for (auto &d : meta.desc) f.line("# Desc: " + d);
    f.line("# meshunit: m");
    f.line("# meshtype: rectangular");
    f.line("# xbase: " + rust_exp(0.5 * g.dx, 17));
    f.line("# ybase: " + rust_exp(0.5 * g.dy, 17));
    f.line("# zbase: " + rust_exp(0.5 * g.dz, 17));
    f.line("# xstepsize: " + rust_exp(g.dx, 17));
    f.line("# ystepsize: " + rust_exp(g.dy, 17));
    f.line("# zstepsize: " + rust_exp(g.dz, 17));
    f.line("# xnodes: " + std::to_string(nx));
    f.line("# ynodes: " + std::to_string(ny));
    f.line("# znodes: " + std::to_string(nz));
    f.line("# xmin: 0");
    f.line("# xmax: " + rust_exp((double)nx * g.dx, 17));
    f.line("# ymin: 0");
    f.line("# ymax: " + rust_exp((double)ny * g.dy, 17));
    f.line("# zmin: 0");
    f.line("# zmax: " + rust_exp(g.dz, 17));
    f.line("# valuedim: 3");
    f.line("# valuelabels: " + meta.labels[0] + " " + meta.labels[1] + " " + meta.labels[2]);
    f.line("# valueunits: " + meta.units[0] + " " + meta.units[1] + " " + meta.units[2]);
    f.line("# End: Header");
    f.line("# Begin: Data Binary 4");
    static_assert(sizeof(float) == 4, "binary4 is IEEE single");
    const float check = 1234567.0f;   // OVF 2.0 binary4 check value (little endian, as x86 / gfx hosts)
    f.bytes(&check, 4);
    std::vector<float> row;
    row.reserve(3 * 4096);
    for (int64_t c = 0; c < nx * ny; ++c) {
        row.push_back((float)m[3 * c]);
        row.push_back((float)m[3 * c + 1]);
        row.push_back((float)m[3 * c + 2]);
        if (row.size() >= 3 * 4096) {
            f.bytes(row.data(), row.size() * 4);
            row.clear();
        }
    }
    if (!row.empty()) f.bytes(row.data(), row.size() * 4);
    f.line("");
    f.line("# End: Data Binary 4");
    f.line("# End: Segment");
    f.line("# End: File");
    f.close();
}

std::string ovf_total_sim_time_line(double t_s) {
    return "Total simulation time:  " + rust_exp(t_s, 16) + "  s";   // OvfMeta::with_total_sim_time (ovf.rs:36-41)
}

// ---------------------------------------------------------------------------
// Host FFT of any length (the kernel cache only)
// ---------------------------------------------------------------------------
namespace {

using C64 = std::complex<double>;

// out[k] = sum_j in[j*stride] w^(jk), w = exp(sign 2 pi i / n): recursive
// decimation in time over the smallest factor (4, 2, 3, 5, then odd primes).
// tw is the table exp(sign 2 pi i t / N) of the top-level length N; a level of
// length n reads it with stride N / n.
void fft_rec(const C64 *in, size_t stride, C64 *out, size_t n, const std::vector<C64> &tw, size_t N,
             std::vector<C64> &scratch) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    size_t p = 0;
    for (size_t q : {size_t(4), size_t(2), size_t(3), size_t(5)})
        if (n % q == 0) { p = q; break; }
    if (!p) {
        for (size_t q = 7; q * q <= n; q += 2)
            if (n % q == 0) { p = q; break; }
        if (!p) p = n;   // prime
    }
    const size_t m = n / p, ts = N / n;
    if (m == 1) {   // direct DFT of a prime length
        for (size_t k = 0; k < n; ++k) {
            C64 s = 0.0;
            for (size_t j = 0; j < n; ++j) s += in[j * stride] * tw[((j * k) % n) * ts];
            out[k] = s;
        }
        return;
    }
    for (size_t r = 0; r < p; ++r) fft_rec(in + r * stride, stride * p, out + r * m, m, tw, N, scratch);
    if (scratch.size() < n) scratch.resize(n);
    C64 t[64];
    std::vector<C64> tbig;
    C64 *tp = t;
    if (p > 64) { tbig.resize(p); tp = tbig.data(); }
    for (size_t k = 0; k < m; ++k) {
        for (size_t r = 0; r < p; ++r) tp[r] = out[r * m + k] * tw[((r * k) % n) * ts];
        for (size_t q = 0; q < p; ++q) {
            C64 s = 0.0;
            for (size_t r = 0; r < p; ++r) s += tp[r] * tw[((r * q) % p) * (N / p)];
            scratch[q * m + k] = s;
        }
    }
    std::copy(scratch.begin(), scratch.begin() + (ptrdiff_t)n, out);
}

std::vector<C64> twiddles(size_t N, int sign) {
    std::vector<C64> tw(N);
    for (size_t t = 0; t < N; ++t) {
        const double a = sign * 2.0 * M_PI * (double)t / (double)N;
        tw[t] = C64(std::cos(a), std::sin(a));
    }
    return tw;
}

// 2D transform of a row-major ny x nx array in place (unnormalised).
void fft2(std::vector<C64> &a, size_t nx, size_t ny, int sign) {
    std::vector<C64> line(std::max(nx, ny)), out(std::max(nx, ny)), scratch;
    const std::vector<C64> twx = twiddles(nx, sign), twy = twiddles(ny, sign);
    for (size_t y = 0; y < ny; ++y) {
        fft_rec(a.data() + y * nx, 1, out.data(), nx, twx, nx, scratch);
        std::copy(out.begin(), out.begin() + (ptrdiff_t)nx, a.begin() + (ptrdiff_t)(y * nx));
    }
    for (size_t x = 0; x < nx; ++x) {
        for (size_t y = 0; y < ny; ++y) line[y] = a[y * nx + x];
        fft_rec(line.data(), 1, out.data(), ny, twy, ny, scratch);
        for (size_t y = 0; y < ny; ++y) a[y * nx + x] = out[y];
    }
}

constexpr char kMagic[8] = {'L', 'L', 'G', 'D', 'M', 'A', 'G', '\0'};
constexpr double kAccuracy = 10.0;   // DEMAG_ACCURACY (demag_fft_uniform.rs:432)

}  // namespace

void kernel_cache_sizes(const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y, int64_t &px, int64_t &py) {
    px = pbc_x > 0 ? g.nx : 2 * g.nx;   // Demag2D::new (demag_fft_uniform.rs:474-478)
    py = pbc_y > 0 ? g.ny : 2 * g.ny;
}

// demag_cache_path (demag_fft_uniform.rs:262-280): Rust `{:.3e}` / `{:.2}`.
std::string kernel_cache_path(const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y) {
    int64_t px, py;
    kernel_cache_sizes(g, pbc_x, pbc_y, px, py);
    char acc[32];
    std::snprintf(acc, sizeof acc, "%.2f", kAccuracy);
    return "out/demag_cache/demag_kernel_nx" + std::to_string(g.nx) + "_ny" + std::to_string(g.ny) +
           "_pbcx" + std::to_string(pbc_x) + "_pbcy" + std::to_string(pbc_y) + "_px" + std::to_string(px) +
           "_py" + std::to_string(py) + "_dx" + rust_exp(g.dx, 3) + "_dy" + rust_exp(g.dy, 3) + "_dz" +
           rust_exp(g.dz, 3) + "_acc" + acc + ".bin";
}

// Real-space kernels at the reference's sizes from the base quadrant
// base[(sy*nx + sx)*4 + c], (sx, sy) >= 0, as the reference's builders lay them
// out: a periodic axis holds the folded bins 0..n-1 directly, an open axis
// holds displacements -(n-1)..n-1 wrapped, with parity (xx, yy, zz even, xy
// odd in each reflected axis) and the Nyquist bin n zero; xy is zero on the
// bin axes.
static void realspace_from_base(const std::vector<double> &base, int64_t nx, int64_t ny, int64_t pbc_x,
                                int64_t pbc_y, int64_t px, int64_t py, std::vector<C64> k[4]) {
    for (int c = 0; c < 4; ++c) k[c].assign((size_t)(px * py), C64(0.0, 0.0));
    for (int64_t iy = 0; iy < py; ++iy) {
        int64_t sy, sgy = 1;
        if (pbc_y > 0) sy = iy;
        else if (iy < ny) sy = iy;
        else if (iy > ny) { sy = py - iy; sgy = -1; }
        else continue;   // Nyquist bin of an open axis
        for (int64_t ix = 0; ix < px; ++ix) {
            int64_t sx, sgx = 1;
            if (pbc_x > 0) sx = ix;
            else if (ix < nx) sx = ix;
            else if (ix > nx) { sx = px - ix; sgx = -1; }
            else continue;
            const double *b = &base[(size_t)((sy * nx + sx) * 4)];
            const size_t o = (size_t)(iy * px + ix);
            k[0][o] = b[0];
            k[1][o] = (ix == 0 || iy == 0) ? 0.0 : (double)(sgx * sgy) * b[1];
            k[2][o] = b[2];
            k[3][o] = b[3];
        }
    }
}

// write_kernel_kspace (demag_fft_uniform.rs:282-320)
void write_kernel_cache(const std::string &path, const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y,
                        const std::vector<double> &base) {
    int64_t px, py;
    kernel_cache_sizes(g, pbc_x, pbc_y, px, py);
    if (px > 0xffffffffLL || py > 0xffffffffLL) throw Error(LLGMI_ERR_INVALID, "grid too large for the cache header");
    std::vector<C64> k[4];
    realspace_from_base(base, g.nx, g.ny, pbc_x, pbc_y, px, py, k);
    for (int c = 0; c < 4; ++c) fft2(k[c], (size_t)px, (size_t)py, -1);
    ensure_parent_dir(path);
    // written beside the target and renamed over it, so a reader never sees a
    // partly written file
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    File f(tmp, "wb");
    f.bytes(kMagic, 8);
    const uint32_t u[7] = {2u, (uint32_t)g.nx, (uint32_t)g.ny, (uint32_t)pbc_x, (uint32_t)pbc_y, (uint32_t)px,
                           (uint32_t)py};
    f.bytes(u, sizeof u);   // little-endian host
    const double d[4] = {g.dx, g.dy, g.dz, kAccuracy};
    f.bytes(d, sizeof d);
    for (int c = 0; c < 4; ++c) f.bytes(k[c].data(), sizeof(C64) * k[c].size());
    f.close();
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw Error(LLGMI_ERR_INVALID, "cannot rename " + tmp + " to " + path);
    }
}

// try_load_kernel_kspace (demag_fft_uniform.rs:322-429): false when the file is
// missing or its header does not match; the base quadrant otherwise.
bool read_kernel_cache(const std::string &path, const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y,
                       std::vector<double> &base) {
    std::FILE *probe = std::fopen(path.c_str(), "rb");
    if (!probe) return false;
    std::fclose(probe);
    int64_t px, py;
    kernel_cache_sizes(g, pbc_x, pbc_y, px, py);
    File f(path, "rb");
    char magic[8];
    f.read(magic, 8);
    uint32_t u[7];
    f.read(u, sizeof u);
    double d[4];
    f.read(d, sizeof d);
    const bool ok = std::memcmp(magic, kMagic, 8) == 0 && u[0] == 2u && u[1] == (uint32_t)g.nx &&
                    u[2] == (uint32_t)g.ny && u[3] == (uint32_t)pbc_x && u[4] == (uint32_t)pbc_y &&
                    u[5] == (uint32_t)px && u[6] == (uint32_t)py && d[0] == g.dx && d[1] == g.dy &&
                    d[2] == g.dz && d[3] == kAccuracy;
    if (!ok) return false;
    std::vector<C64> k[4];
    for (int c = 0; c < 4; ++c) {
        k[c].resize((size_t)(px * py));
        f.read(k[c].data(), sizeof(C64) * k[c].size());
    }
    const double scale = 1.0 / (double)(px * py);
    base.assign((size_t)(4 * g.nx * g.ny), 0.0);
    for (int c = 0; c < 4; ++c) {
        fft2(k[c], (size_t)px, (size_t)py, +1);
        for (int64_t sy = 0; sy < g.ny; ++sy)
            for (int64_t sx = 0; sx < g.nx; ++sx)
                base[(size_t)((sy * g.nx + sx) * 4 + c)] = k[c][(size_t)(sy * px + sx)].real() * scale;
    }
    return true;
}

namespace {
constexpr char kDkMagic[8] = {'L', 'L', 'G', 'D', 'K', 'H', '8', '\0'};
constexpr size_t kDkFields = sizeof(DkCacheKey) / sizeof(uint64_t);
static_assert(kDkFields == 23, "DeltaKernelKey has 23 fields");
std::string hex16(uint64_t v) {
    char b[24];
    std::snprintf(b, sizeof b, "%016llx", (unsigned long long)v);
    return b;
}
double from_bits(uint64_t u) {
    double d;
    std::memcpy(&d, &u, 8);
    return d;
}
}  // namespace

std::string dk_cache_path(const DkCacheKey &k) {
    // MGOperatorSettings::tag (:806-827)
    std::string st = k.stencil == 0 ? "7" : k.stencil == 1 ? "iso9"
                   : (k.iso27_alpha_bits != 0 ? "iso27a" + hex16(k.iso27_alpha_bits) : "iso27");
    const std::string op = st + "_" + (k.prolong == 0 ? "inj" : "tri") + "_" + (k.coarse_op == 0 ? "rd" : "gal");
    const char *bc = k.bc == 0 ? "dir0" : k.bc == 1 ? "dip" : "bh";
    const char *sm = k.smoother == 0 ? "wj" : "rb";
    return "out/demag_cache/demag_mg_hybrid_dk_nx" + std::to_string(k.nx) + "_ny" + std::to_string(k.ny) + "_dx" +
           rust_exp(from_bits(k.dx_bits), 3) + "_dy" + rust_exp(from_bits(k.dy_bits), 3) + "_dz" +
           rust_exp(from_bits(k.dz_bits), 3) + "_pad" + std::to_string(k.pad_xy) + "_nvacz" +
           std::to_string(k.n_vac_z) + "_op" + op + "_bc" + bc + "_th" + rust_exp(from_bits(k.tree_theta_bits), 3) +
           "_leaf" + std::to_string(k.tree_leaf) + "_dep" + std::to_string(k.tree_max_depth) + "_sm" + sm + "_pre" +
           std::to_string(k.pre_smooth) + "_post" + std::to_string(k.post_smooth) + "_om" + hex16(k.omega_bits) +
           "_som" + hex16(k.sor_omega_bits) + "_sig" + hex16(k.sigma_bits) + "_r" + std::to_string(k.radius_xy) +
           "_dv" + std::to_string(k.delta_v_cycles) + ".bin";
}

// save_delta_kernel_to_disk (:2966-3052)
void write_dk_cache(const std::string &path, const DkCacheKey &k, const std::vector<double> &dk4) {
    const size_t n = (size_t)((2 * k.radius_xy + 1) * (2 * k.radius_xy + 1));
    if (dk4.size() != 4 * n) throw Error(LLGMI_ERR_INVALID, "correction stencil size does not match its radius");
    ensure_parent_dir(path);
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    File f(tmp, "wb");
    f.bytes(kDkMagic, 8);
    f.bytes(&k, sizeof k);   // 23 little-endian u64 in the reference's order
    f.bytes(dk4.data(), sizeof(double) * dk4.size());
    const std::vector<double> zeros(3 * n, 0.0);
    f.bytes(zeros.data(), sizeof(double) * zeros.size());
    f.close();
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw Error(LLGMI_ERR_INVALID, "cannot rename " + tmp + " to " + path);
    }
}

// load_delta_kernel_from_disk (:2821-2964); the PPPM-phi stencils that follow
// are not read
bool read_dk_cache(const std::string &path, const DkCacheKey &k, std::vector<double> &dk4) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    uint64_t u[kDkFields];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kDkMagic, 8) == 0 &&
              std::fread(u, sizeof(uint64_t), kDkFields, f) == kDkFields && std::memcmp(u, &k, sizeof k) == 0;
    if (ok) {
        const size_t n = (size_t)((2 * k.radius_xy + 1) * (2 * k.radius_xy + 1));
        std::vector<double> v(4 * n);
        ok = std::fread(v.data(), sizeof(double), v.size(), f) == v.size();
        if (ok) dk4.swap(v);
    }
    std::fclose(f);
    return ok;
}

}  // namespace llgmi
