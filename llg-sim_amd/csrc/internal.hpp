// Host-side declarations shared by the translation units of libllgmi.so.
#pragma once

#include "common.hpp"

#include <vector>

namespace llgmi {

// ---------------------------------------------------------------------------
// Per-kernel-class timing with HIP events recorded on the launch stream.
// ---------------------------------------------------------------------------
struct Profiler {
    unsigned mask = 0;  // bit c set: time kernel class c
    struct Pending {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    static constexpr int kClasses = 16;   // >= LLGMI_K_COUNT
    double total_ms[kClasses] = {0};
    int64_t count[kClasses] = {0};

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        LLGMI_HIP(hipEventCreate(&e));
        return e;
    }
    void drain() {
        for (auto &p : pending) {
            LLGMI_HIP(hipEventSynchronize(p.b));
            float ms = 0.f;
            LLGMI_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            total_ms[p.cls] += ms;
            count[p.cls] += 1;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() {
        drain();
        for (int i = 0; i < kClasses; ++i) { total_ms[i] = 0; count[i] = 0; }
    }
    ~Profiler() {
        for (auto &p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

struct ProfScope {
    Profiler *p;
    int cls;
    hipStream_t s;
    hipEvent_t a{}, b{};
    ProfScope(Profiler *p_, int c, hipStream_t s_) : p(p_), cls(c), s(s_) {
        if (p && (p->mask >> cls) & 1u) {
            if (p->pending.size() > 4096) p->drain();
            a = p->get();
            b = p->get();
            LLGMI_HIP(hipEventRecord(a, s));
        }
    }
    ~ProfScope() {
        if (p && (p->mask >> cls) & 1u) {
            if (hipEventRecord(b, s) == hipSuccess) p->pending.push_back({cls, a, b});
        }
    }
};

// ---------------------------------------------------------------------------
// FFT demag operator for one grid (Ms independent), precision T.
// ---------------------------------------------------------------------------
template <typename T>
struct LocalArgs;

// rq_col flag: the column pass runs as persistent blocks (k_col_demag PIPE)
constexpr int kColPipe = 8;

template <typename T>
struct DemagFFT {
    bool built = false;
    int64_t nx = 0, ny = 0, px = 0, py = 0, hx = 0, hy = 0;
    int64_t pbc_x = 0, pbc_y = 0;   // periodic image counts (0 = open axis)
    int lx = 0, ly = 0;
    double dx = 0, dy = 0, dz = 0;
    cpx<T> *tw_x = nullptr, *tw_y = nullptr;
    T *kern = nullptr;  // 4 planes [Kxx, Kxy, Kyy, Kzz], each [hx][hy]
    cpx<T> *spec = nullptr;    // [3][hx][ny]
    int rq_row = 4, rq_col = 4, rq_row_inv = 4;  // log2 first/last radix, picked by autotune()
    int rq_fused = 0;   // inverse row pass fused with the local field: log2 radix, -1 = unfused, 0 = untuned
    // real-space base quadrant of the kernel, [(sy*nx + sx)*4 + {xx, xy, yy, zz}]
    // (periodic axes: folded bins), kept on the host for the kernel cache
    std::vector<double> base_host;

    ~DemagFFT();
    void release();
    // pbc_x / pbc_y > 0: periodic axis with MuMax-style image sums; the
    // convolution stays zero padded (a periodic kernel folded onto the
    // displacements -(n-1)..n-1 gives the circular convolution of period n).
    // base_in: a base quadrant from the kernel cache instead of the quadrature.
    void build(int64_t nx, int64_t ny, double dx, double dy, double dz, hipStream_t stream,
               int64_t pbc_x = 0, int64_t pbc_y = 0, const std::vector<double> *base_in = nullptr);
    // bd = B_demag (overwrite) for M = ms * m; m and bd are SoA with 3*nx*ny entries.
    void field(const T *m, T ms, T *bd, hipStream_t stream, Profiler *prof);
    // Times the radix-8 and radix-16 variants of each pass on this grid and keeps the faster.
    void autotune(hipStream_t stream, int forced);
    void launch_row_fwd(const T *m, T ms, hipStream_t stream, int rq);
    void launch_col(hipStream_t stream, int rq);
    void launch_row_inv(T *bd, hipStream_t stream, int rq);
    // b (overwrite) = local field terms + B_demag of M = ms * m, with the inverse
    // row FFT and the local field in one kernel (kout: torque, or null);
    // bd_scratch (3 n) serves the unfused variant
    void field_local(const T *m, T ms, const LocalArgs<T> &a, const uint8_t *geom, T *b, T *kout,
                     T *bd_scratch, hipStream_t stream, Profiler *prof);
    void launch_row_inv_local(const LocalArgs<T> &a, const T *m, const uint8_t *geom, T *b, T *kout,
                              hipStream_t stream, int rq);
    void read_kernel(std::vector<double> &out) const;
};

// ---------------------------------------------------------------------------
// Local field kernel arguments (all coefficients precomputed in FP64).
// ---------------------------------------------------------------------------
template <typename T>
struct LocalArgs {
    int nx, ny;
    T bext[3];
    int exch_on;           // A != 0 && Ms != 0
    T exch_coeff;          // 2A/Ms
    T inv_dx2, inv_dy2;
    int has_eta;           // DMI != 0 (chiral ghosts in exchange)
    T eta, dx, dy;
    int anis_on;
    T anis_coeff;          // 2Ku/Ms
    T u[3];
    int dmi_on;            // mask includes DMI and D != 0, A != 0, Ms != 0
    T dmi_pref;            // 2D/Ms
    T inv_2dx, inv_2dy;
    int torque;            // also write k: 1 = llg_rhs(m, B), 2 = llg_rhs_relax(m, B)
    T rhs_pref;            // -gamma/(1+alpha^2), or -gamma alpha/(1+alpha^2) for relax
    T alpha;
};

template <typename T>
void launch_local_field(const LocalArgs<T> &a, const T *m, const T *bd, const uint8_t *geom,
                        T *b, T *k, hipStream_t s);

template <typename T>
void launch_aos_to_soa(const double *aos, T *soa, int64_t n, hipStream_t s);
template <typename T>
void launch_soa_to_aos(const T *soa, double *aos, int64_t n, hipStream_t s);
template <typename T>
void launch_soa_add_to_aos(const T *soa, double *aos, int64_t n, hipStream_t s);
// dst[i] += src[i], i < count
template <typename T>
void launch_soa_accumulate(const T *src, T *dst, int64_t count, hipStream_t s);
template <typename T>
void launch_mask_cells(T *m, const uint8_t *keep, int64_t n, hipStream_t s);

// RK45 stage state: out = normalize(m + sum_q c[q] k[q]) applied in order q = 0..nk-1.
template <typename T>
struct StageArgs {
    const T *k[6];
    T c[6];
    int nk;
    // with dt non-null the coefficients are T(*dt * a[q]) (the same double
    // product the host forms for c[q]), so a captured launch follows the step size
    const double *dt;
    double a[6];
};
template <typename T>
void launch_rk_stage(const T *m, const StageArgs<T> &a, T *out, int64_t n, hipStream_t s);

// Error norm max_{cell,comp} |tau5 - tau4| over k1,k3,k4,k5,k6,k7 (FP64 arithmetic).
// partial must hold kErrBlocks doubles; result written to out[0].
static constexpr int kErrBlocks = 1024;
template <typename T>
void launch_rk_err(const T *k1, const T *k3, const T *k4, const T *k5, const T *k6,
                   const T *k7, int64_t n, bool combined, double *partial, double *out,
                   hipStream_t s);

// ---------------------------------------------------------------------------
// Relaxation path (relax.hip)
// ---------------------------------------------------------------------------
static constexpr int kRelaxRedBlocks = 1024;   // partial rows of the reductions below

// dt-scaled Bogacki-Shampine weights: b = third order (k1..k3), bh = embedded
// second order (k1..k4)
struct Rk23ErrCoef {
    double b[3];
    double bh[4];
};
// eps = max |m4 - m2nd| (llg.rs:399-420), written to out[0]; partial holds
// kRelaxRedBlocks doubles.
template <typename T>
void launch_rk23_err(const T *m, const T *k1, const T *k2, const T *k3, const T *k4, const T *m4,
                     const Rk23ErrCoef &c, int64_t n, double *partial, double *out, hipStream_t s);

// compute_energy_geom (energy.rs:80-283) coefficients
struct EnergyArgs {
    int nx, ny;
    double dx, dy, v;      // v = cell volume
    double aex, ku, ms;
    double u[3], bext[3];
    int dmi_on;            // Some(D != 0) and A != 0
    double pref, eta;      // 2D/Ms, D/(2A)
};
// out[0..4] = exchange, anisotropy, zeeman, dmi, demag; partial: 5 kRelaxRedBlocks doubles;
// bd (demag field) and geom may be null.
template <typename T>
void launch_energy(const EnergyArgs &a, const T *m, const T *bd, const uint8_t *geom,
                   double *partial, double *out, hipStream_t s);
// out[0..2] = max |m x B|, sum |m x B|, sum |m x B|^2; partial: 3 kRelaxRedBlocks doubles
template <typename T>
void launch_torque(const T *m, const T *b, int64_t n, double *partial, double *out, hipStream_t s);

}  // namespace llgmi
