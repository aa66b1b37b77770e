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
    int lx = 0, ly = 0;   // length codes of px and py: log2 for powers of two, 32 + log2(M) for 3M
    double dx = 0, dy = 0, dz = 0;
    cpx<T> *tw_x = nullptr, *tw_y = nullptr;
    T *kern = nullptr;  // 4 planes [Kxx, Kxy, Kyy, Kzz], each [hx][hy], times 1 / (px py)
    cpx<T> *spec = nullptr;    // [3][hx][ny]
    // full-length forward column spectra [3][px/2][py] of the split column pass
    // (columns whose paired block would not fit in LDS), else null
    cpx<T> *colfull = nullptr;
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
    // columns taken by the split column pass (see demag_fft.hip)
    bool col_split() const;
    // a length is 3 * 2^k along an axis (such operators are not in the batched kernels)
    bool has_f3() const { return lx >= 32 || ly >= 32; }
    void launch_col(hipStream_t stream, int rq);
    void launch_row_inv(T *bd, hipStream_t stream, int rq);
    // b (overwrite) = local field terms + B_demag of M = ms * m, with the inverse
    // row FFT and the local field in one kernel (kout: torque, or null);
    // bd_scratch (3 n) serves the unfused variant
    void field_local(const T *m, T ms, const LocalArgs<T> &a, const uint8_t *geom, T *b, T *kout,
                     T *bd_scratch, hipStream_t stream, Profiler *prof);
    void launch_row_inv_local(const LocalArgs<T> &a, const T *m, const uint8_t *geom, T *b, T *kout,
                              hipStream_t stream, int rq);
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
    // Next RK stage state, formed after the torque (st_out non-null, torque
    // written): st_out = normalize(st_m + sum_q c_q k_q), q = 0 .. st_nk - 1 in
    // order, with k_q = st_k[q] read from memory for q < st_nk - 1 and the
    // cell's own torque for the last q; c_q = T(*st_dt * st_a[q]) when st_dt is
    // set, else st_c[q].  The same arithmetic as the stand-alone stage kernel.
    T *st_out;
    const T *st_m;
    const T *st_k[5];
    T st_c[6];
    const double *st_dt;
    double st_a[6];
    int st_nk;
};

template <typename T>
void launch_local_field(const LocalArgs<T> &a, const T *m, const T *bd, const uint8_t *geom,
                        T *b, T *k, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched films: one record per film in device memory; a launch covers all
// films of one kernel class through a table of (film, block) entries.
// ---------------------------------------------------------------------------
template <typename T>
struct BFilm {
    // FFT demag operator of the film (fft = 0: no demag term)
    cpx<T> *spec;
    const cpx<T> *tw_x, *tw_y;
    const T *kern;
    int nx, ny, px, py;
    T ms;
    int fft;
    int skip1;           // nonzero: the stage-1 field evaluation is skipped (FSAL)
    int accept;          // the controller's decision (for the accepted-state copy)
    int e0, ne;          // this film's entries in the element table
    double dt;           // step size of the attempt
    LocalArgs<T> la;     // local terms and torque (llg_rhs)
    const T *in[3];      // field-evaluation inputs: m, stage state, m7
    T *m, *stage, *m7, *b, *bd;
    T *k[8];             // k1..k7 (k[0] unused)
};
// an entry of a launch table: film index and the block's index within the film
struct BEntry {
    int film, blk;
};

// Field evaluation of every film whose entries the tables hold (FFT passes by
// kernel class, see DemagFFT), element-wise kernels over all films:
// which = input (0 m, 1 stage state, 2 m7), ks = torque output k[ks];
// stage1 = skip films whose skip1 is set.
template <typename T>
void launch_batch_local(const BFilm<T> *F, const BEntry *tab, int nent, int demag, int which, int ks, int stage1,
                        hipStream_t s);
template <typename T>
void launch_batch_stage(const BFilm<T> *F, const BEntry *tab, int nent, int stage, hipStream_t s);
// error norm of every film: partial[e] per element-table entry, err[f] per film
template <typename T>
void launch_batch_err(const BFilm<T> *F, const BEntry *tab, int nent, int nfilms, double *partial, double *err,
                      hipStream_t s);
// m = m7 for the films whose accept is set
template <typename T>
void launch_batch_accept(const BFilm<T> *F, const BEntry *tab, int nent, hipStream_t s);
// m = m7 and k1 = k7 when *flag is set (3 n values each).
template <typename T>
void launch_rk_accept(const int *flag, T *m, const T *m7, T *k1, const T *k7, int64_t n, hipStream_t s);
// The RK45 controller per film on the device for attempt a (outputs at
// [a * nfilms + film]); updates dt, accept, FSAL skip and the k1/k7 pointers
// of each film record.
template <typename T>
void launch_batch_control(BFilm<T> *F, int nfilms, const double *err, double max_err, double headroom, double dt_min,
                          double dt_max, int a, double *o_eps, int *o_acc, double *o_dtu, hipStream_t s);
// cells per element-table entry
constexpr int kBatchCells = 256;

// Batched FFT passes over the films of one kernel class (the same lengths and
// radix choices): kind 0 row R2C of in[which] (entries comp * row blocks +
// row block), 1 column pass (entries = items), 2 row C2R into bd (as 0),
// 3 row C2R fused with the local field into b and k[ks] (entries = row
// blocks).  lg: log2 length of the pass, rq: its first/last radix, half: the
// film's nx (rows) or ny (columns) is half the length.
template <typename T>
void launch_batch_fft(int kind, int lg, int rq, bool half, const BFilm<T> *F, const BEntry *tab, int nent,
                      int which, int ks, int stage1, hipStream_t s);
// row blocks of a film's row pass (kind 0 / 2) or fused inverse (kind 3);
// for the column pass (kind 1, ny = item count) the blocks of items
template <typename T>
int batch_row_blocks(int kind, int lg, int rq, int ny);
// The representative length of the launch class of (kind, lg, rq): one
// launch covers the films of every length whose blocks have the same shape.
template <typename T>
int batch_class_lg(int kind, int lg, int rq);
// Opts the class's kernels into the LDS they need.
template <typename T>
void batch_allow_lds(int kind, int lg, int rq);

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
// Stage 2 inside a device-controlled run: with *flag set (the previous attempt
// accepted) first m = m7 and k1 = k7, then out = normalize(m + c k1).
template <typename T>
void launch_rk_stage2_accept(const int *flag, T *m, const T *m7, T *k1, const T *k7, const StageArgs<T> &a,
                             T *out, int64_t n, hipStream_t s);

// The RK45 controller of one film on the device for attempt a: outputs at
// [a], the next step size into *dt and the decision into *flag.
struct RkControl {
    double *dt = nullptr;
    int *flag = nullptr;
    double max_err = 0, headroom = 0, dt_min = 0, dt_max = 0;
    int a = 0;
    double *o_eps = nullptr;
    int *o_acc = nullptr;
    double *o_dtu = nullptr;
};

// Error norm max_{cell,comp} |tau5 - tau4| over k1,k3,k4,k5,k6,k7 (FP64 arithmetic).
// partial must hold kErrBlocks doubles; result written to out[0].  With ctl,
// the kernel that forms the maximum then runs the controller on it.
static constexpr int kErrBlocks = 1024;
template <typename T>
void launch_rk_err(const T *k1, const T *k3, const T *k4, const T *k5, const T *k6,
                   const T *k7, int64_t n, bool combined, double *partial, double *out,
                   hipStream_t s, const RkControl *ctl = nullptr);

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
