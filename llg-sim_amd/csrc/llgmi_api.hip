// C ABI of libllgmi.so (include/llgmi.h): contexts, argument checks, dispatch
// to the per-precision engine, and the RK45 driver with its FP64 controller.
#include "../../include/llgmi.h"
#include "common.hpp"
#include "internal.hpp"
#include "formats.hpp"
#include "mg.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

using namespace llgmi;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace

// the other translation units report errors through the same last-error slot
int llgmi::set_last_error(int code, const std::string &msg) { return fail(code, msg); }

namespace {

// LLG_DEMAG_METHOD, parsed once (demag.rs:35-44, params.rs:76-83).
int env_demag_method() {
    static int v = [] {
        const char *s = std::getenv("LLG_DEMAG_METHOD");
        if (!s) return -1;
        std::string t(s);
        auto b = t.find_first_not_of(" \t\r\n");
        auto e = t.find_last_not_of(" \t\r\n");
        t = b == std::string::npos ? std::string() : t.substr(b, e - b + 1);
        for (auto &ch : t) ch = (char)std::tolower((unsigned char)ch);
        if (t == "fft" || t == "fftuniform" || t == "uniform") return LLGMI_DEMAG_FFT;
        if (t == "mg" || t == "multigrid" || t == "poissonmg" || t == "poisson_mg") return LLGMI_DEMAG_MG;
        if (t == "dst" || t == "poissondst" || t == "poisson_dst" || t == "dst_open") return LLGMI_DEMAG_DST;
        return -1;
    }();
    return v;
}

int resolved_demag_method(const llgmi_material &mat) {
    int e = env_demag_method();
    return e >= 0 ? e : mat.demag_method;
}

struct EngineBase {
    virtual ~EngineBase() = default;
};

}  // namespace

// ---------------------------------------------------------------------------
// Engine: all device state of a context, in precision T.
// ---------------------------------------------------------------------------
template <typename T>
struct Engine : EngineBase {
    llgmi_grid g;
    int64_t n;
    hipStream_t stream;
    Profiler *prof;

    T *d_m = nullptr, *d_b = nullptr, *d_bd = nullptr;
    double *d_aos = nullptr;
    uint8_t *d_geom = nullptr;
    std::vector<double> h_tmp;

    DemagFFT<T> fft;
    DemagFFT<T> fft_pbc;   // the last periodic operator (keyed by pbc_x, pbc_y)
    std::unique_ptr<MGSolver> mg;
    bool has_mg_cfg = false;       // llgmi_ctx_set_mg_config; otherwise the environment per call
    llgmi_mg_config mg_cfg{};
    bool has_hyb_cfg = false;      // llgmi_ctx_set_mg_hybrid_config; otherwise the environment per call
    llgmi_mg_hybrid_config hyb_cfg{};
    // hybrid correction stencil (DeltaKernel2D) and the settings it was built for
    std::vector<double> dk_host;
    double *d_dk = nullptr, *d_hyb_scale = nullptr;
    std::string dk_key;

    // RK45 scratch (RK45Scratch, llg.rs:126-189): k1..k7, stage state, m7.
    T *d_k[8] = {nullptr};
    T *d_stage = nullptr, *d_stage2 = nullptr, *d_m7 = nullptr;
    double *d_partial = nullptr, *d_err = nullptr;
    double *h_err = nullptr;  // pinned
    bool fsal_valid = false;
    int last_mask = -1;
    int issued_mask = -1;    // mask of the attempt rk45_issue left for rk45_finish
    // HIP-graph replay of RK45 attempts: one graph per (FSAL used, k1/k7
    // swapped), keyed by what the captured launches read from the host;
    // `seen` is the configuration run eagerly once before it is captured
    struct RkGraph {
        hipGraphExec_t exec = nullptr;
        std::string key, seen;
    };
    RkGraph rk_graph[2][2];
    // device-controlled runs (rk45_run): the last captured run and its
    // per-attempt outputs (device and pinned host memory)
    RkGraph run_graph;
    double *d_run_eps = nullptr, *d_run_dtu = nullptr, *h_run_eps = nullptr, *h_run_dtu = nullptr;
    int *d_run_acc = nullptr, *h_run_acc = nullptr, *d_run_flag = nullptr;
    int64_t run_cap = 0;
    bool k_swapped = false;
    uint64_t fft_gen = 0;                      // bumped whenever an FFT operator is (re)built
    double *d_dt = nullptr, *h_dt = nullptr;   // the attempt's step size (pinned host -> device)
    double last_bext[3] = {0, 0, 0};

    // RK23 relax scratch (RK23Scratch, llg.rs:196-268): k1..k4, b4 and m4 of
    // the last step, FSAL / last-field state.  Stage states 2 and 3 share d_stage.
    T *d_r23k[5] = {nullptr};
    T *d_r23b4 = nullptr, *d_r23m4 = nullptr;
    bool r23_b4_valid = false;   // last_b_eff_valid
    bool r23_has_last = false;   // last_mask / last_b_ext are Some
    int r23_last_mask = -1;
    double r23_last_bext[3] = {0, 0, 0};
    // reduction rows and results of the energy / torque / RK23 error kernels
    double *d_red = nullptr, *d_redout = nullptr, *h_red = nullptr;

    Engine(const llgmi_grid &g_, hipStream_t s, Profiler *p) : g(g_), n(g_.nx * g_.ny), stream(s), prof(p) {}

    ~Engine() override {
        void *ps[] = {d_m, d_b, d_bd, d_aos, d_geom, d_stage, d_stage2, d_m7, d_partial, d_err,
                      d_r23b4, d_r23m4, d_red, d_redout, d_dk, d_hyb_scale};
        for (void *q : ps)
            if (q) (void)hipFree(q);
        for (int i = 0; i < 8; ++i)
            if (d_k[i]) (void)hipFree(d_k[i]);
        for (int i = 0; i < 5; ++i)
            if (d_r23k[i]) (void)hipFree(d_r23k[i]);
        if (h_err) (void)hipHostFree(h_err);
        if (h_red) (void)hipHostFree(h_red);
        if (h_dt) (void)hipHostFree(h_dt);
        if (d_dt) (void)hipFree(d_dt);
        drop_rk_graphs();
        free_run_outputs();
        if (d_run_flag) (void)hipFree(d_run_flag);
    }
    void free_run_outputs() {
        void *dv[] = {d_run_eps, d_run_dtu, d_run_acc};
        for (void *q : dv)
            if (q) (void)hipFree(q);
        void *hv[] = {h_run_eps, h_run_dtu, h_run_acc};
        for (void *q : hv)
            if (q) (void)hipHostFree(q);
        d_run_eps = d_run_dtu = h_run_eps = h_run_dtu = nullptr;
        d_run_acc = h_run_acc = nullptr;
        run_cap = 0;
    }

    template <typename U>
    void alloc(U *&p, size_t count) {
        if (!p) LLGMI_HIP(hipMalloc(&p, sizeof(U) * std::max<size_t>(count, 1)));
    }

    void ensure_io() {
        alloc(d_m, 3 * n);
        alloc(d_b, 3 * n);
        alloc(d_aos, 3 * n);
    }

    // LLGMI_DEMAG_CACHE=1: load / save the reference's kernel cache file around
    // each operator build (Demag2D::new, demag_fft_uniform.rs:497-598); write
    // failures only warn, as there.
    static bool auto_cache() {
        const char *e = std::getenv("LLGMI_DEMAG_CACHE");
        return e && std::string(e) == "1";
    }
    // Every (re)build of an operator frees the buffers that captured RK45
    // attempts name: the generation in the graph key changes first (so a
    // failed build cannot leave a stale key behind) and the captured graphs
    // are dropped.
    void rebuild_fft(DemagFFT<T> &op, int64_t pbc_x, int64_t pbc_y, const std::vector<double> *base) {
        ++fft_gen;
        drop_rk_graphs();
        op.build(g.nx, g.ny, g.dx, g.dy, g.dz, stream, pbc_x, pbc_y, base);
    }
    void drop_rk_graphs() {
        for (auto &row : rk_graph)
            for (auto &G : row) {
                if (G.exec) (void)hipGraphExecDestroy(G.exec);
                G = RkGraph{};
            }
        if (run_graph.exec) (void)hipGraphExecDestroy(run_graph.exec);
        run_graph = RkGraph{};
    }
    void build_fft(DemagFFT<T> &op, int64_t pbc_x, int64_t pbc_y) {
        std::vector<double> base;
        const std::string path = kernel_cache_path(g, pbc_x, pbc_y);
        bool cached = false;
        if (auto_cache()) {
            // as try_load_kernel_kspace(..).unwrap_or(false): an unreadable or
            // short file is a miss, rebuilt and written again
            try {
                cached = read_kernel_cache(path, g, pbc_x, pbc_y, base);
            } catch (const std::exception &e) {
                std::fprintf(stderr, "[llgmi] WARN: demag kernel cache not read (%s)\n", e.what());
                cached = false;
            }
        }
        rebuild_fft(op, pbc_x, pbc_y, cached ? &base : nullptr);
        if (auto_cache() && !cached) {
            try {
                write_kernel_cache(path, g, pbc_x, pbc_y, op.base_host);
            } catch (const std::exception &e) {
                std::fprintf(stderr, "[llgmi] WARN: demag kernel cache not written (%s)\n", e.what());
            }
        }
    }
    void ensure_fft() {
        if (!fft.built) build_fft(fft, 0, 0);
    }
    // the operator for (pbc_x, pbc_y), built if needed
    DemagFFT<T> &fft_for(int64_t pbc_x, int64_t pbc_y) {
        if (pbc_x == 0 && pbc_y == 0) {
            ensure_fft();
            return fft;
        }
        if (pbc_x > (1 << 20) || pbc_y > (1 << 20))
            throw Error(LLGMI_ERR_UNSUPPORTED, "periodic image count beyond 2^20");
        if (!fft_pbc.built || fft_pbc.pbc_x != pbc_x || fft_pbc.pbc_y != pbc_y) build_fft(fft_pbc, pbc_x, pbc_y);
        return fft_pbc;
    }

    // add_demag_field_poisson_mg (demag_poisson_mg.rs:3612-3632): configuration
    // per call, cached solver rebuilt on a structure change, runtime knobs applied.
    MGSolver &ensure_mg() {
        llgmi_mg_config c;
        if (has_mg_cfg) c = mg_cfg;
        else mg_config_from_env(c);
        if (mg && !mg->same_structure(g, c)) mg.reset();
        if (!mg) mg.reset(new MGSolver(g, c, stream, prof));
        mg->apply_cfg(c);
        return *mg;
    }

    // B_demag into d_bd (overwrite).  method resolved by the caller.
    void demag_into_bd(const T *m, const llgmi_material &mat, int method, uint64_t pbc_x,
                       uint64_t pbc_y) {
        alloc(d_bd, 3 * n);
        // demag.rs:109-141: MG and DST fall back to FFT with periodic axes
        if (pbc_x > 0 || pbc_y > 0) {
            if (pbc_x > (uint64_t)(1 << 20) || pbc_y > (uint64_t)(1 << 20))
                throw Error(LLGMI_ERR_UNSUPPORTED, "periodic image count beyond 2^20");
            fft_for((int64_t)pbc_x, (int64_t)pbc_y).field(m, T(mat.ms), d_bd, stream, prof);
        } else if (method == LLGMI_DEMAG_FFT) {
            ensure_fft();
            fft.field(m, T(mat.ms), d_bd, stream, prof);
        } else {
            MGSolver &mg = ensure_mg();
            llgmi_mg_hybrid_config h;
            if (has_hyb_cfg) h = hyb_cfg;
            else mg_hybrid_config_from_env(h);
            if (!(h.enabled && h.radius_xy > 0)) {
                mg.template demag_field<T>(m, mat.ms, d_bd);
                return;
            }
            // DemagPoissonMGHybrid::add_field (demag_poisson_mg.rs:2512-2528)
            const int64_t r = std::min({h.radius_xy, std::max<int64_t>(g.nx - 2, 0) / 2,
                                        std::max<int64_t>(g.ny - 2, 0) / 2});
            if (r > 0) ensure_delta_kernel(h, r, mg.config(), mat);
            mg.template demag_field<T>(m, mat.ms, d_bd, h.sigma_cells, true);
            if (r > 0) {
                if (h.bscale > 0.0) alloc(d_hyb_scale, n);
                launch_hybrid_correction<T>(d_dk, (int)r, m, d_bd, mat.ms, h.bscale, (int)h.bmargin, d_hyb_scale,
                                            (int)g.nx, (int)g.ny, stream);
            }
        }
    }

    // ensure_delta_kernel (demag_poisson_mg.rs:2412-2466): rebuilt when the MG
    // settings, the hybrid settings or the effective radius change.  With
    // LLG_DEMAG_MG_HYBRID_CACHE=1 the stencil is read from, or after a build
    // written to, the reference's LLGDKH8 file; a missing, short or foreign
    // file is a miss, and a failed write only costs the next process a
    // rebuild (as in the reference).
    void ensure_delta_kernel(const llgmi_mg_hybrid_config &h, int64_t r, const llgmi_mg_config &c,
                             const llgmi_material &mat) {
        llgmi_mg_hybrid_config hk = h;
        hk.bscale = 0.0;
        hk.bmargin = 0;
        std::string key((const char *)&c, sizeof c);
        key.append((const char *)&hk, sizeof hk);
        key.append((const char *)&r, sizeof r);
        if (d_dk && key == dk_key) return;
        const char *ce = std::getenv("LLG_DEMAG_MG_HYBRID_CACHE");
        const bool disk = ce && std::string(ce) == "1";
        const DkCacheKey dkey = dk_cache_key(h, r, c);
        if (disk) {
            std::vector<double> v;
            if (read_dk_cache(dk_cache_path(dkey), dkey, v)) {
                dk_host.swap(v);
                upload_dk();
                dk_key = key;
                return;
            }
        }
        build_delta_kernel(h, r, c, mat);
        dk_key = key;
        if (disk) {
            try {
                write_dk_cache(dk_cache_path(dkey), dkey, dk_host);
            } catch (const Error &) {
            }
        }
    }
    // DeltaKernelKey::new (demag_poisson_mg.rs:2305-2333) for radius r
    DkCacheKey dk_cache_key(const llgmi_mg_hybrid_config &h, int64_t r, const llgmi_mg_config &c) const {
        auto bits = [](double d) {
            uint64_t u;
            std::memcpy(&u, &d, 8);
            return u;
        };
        DkCacheKey k{};
        k.nx = (uint64_t)g.nx;
        k.ny = (uint64_t)g.ny;
        k.dx_bits = bits(g.dx);
        k.dy_bits = bits(g.dy);
        k.dz_bits = bits(g.dz);
        k.pad_xy = (uint64_t)c.pad_xy;
        k.n_vac_z = (uint64_t)c.n_vac_z;
        k.stencil = (uint64_t)c.stencil;
        k.prolong = (uint64_t)c.prolong;
        k.coarse_op = (uint64_t)c.coarse_op;
        k.iso27_alpha_bits = c.iso27_alpha > 0.0 ? bits(c.iso27_alpha) : 0;
        k.bc = (uint64_t)c.bc;
        k.tree_theta_bits = bits(c.tree_theta);
        k.tree_leaf = (uint64_t)c.tree_leaf;
        k.tree_max_depth = (uint64_t)c.tree_max_depth;
        k.pre_smooth = (uint64_t)c.pre_smooth;
        k.post_smooth = (uint64_t)c.post_smooth;
        k.smoother = (uint64_t)c.smoother;
        k.omega_bits = bits(c.omega);
        k.sor_omega_bits = bits(c.sor_omega);
        k.sigma_bits = bits(h.sigma_cells);
        k.radius_xy = (uint64_t)r;
        k.delta_v_cycles = (uint64_t)h.delta_v_cycles;
        return k;
    }
    void upload_dk() {
        if (d_dk) (void)hipFree(d_dk);
        d_dk = nullptr;
        LLGMI_HIP(hipMalloc(&d_dk, sizeof(double) * dk_host.size()));
        LLGMI_HIP(hipMemcpy(d_dk, dk_host.data(), sizeof(double) * dk_host.size(), hipMemcpyHostToDevice));
    }

    // build_delta_kernel_impulse (demag_poisson_mg.rs:3159-3597), without its
    // diagnostic branches: impulse responses of the FFT and of a fixed-cycle MG
    // solve at the film centre, dK = (B_fft - B_mg) / Ms in the window, then the
    // reference's in-place symmetrisation and uniform-bias fix on the host.
    void build_delta_kernel(const llgmi_mg_hybrid_config &h, int64_t r, const llgmi_mg_config &c,
                            const llgmi_material &mat) {
        const int64_t nx = g.nx, ny = g.ny, cx = nx / 2, cy = ny / 2, st = 2 * r + 1, nst = st * st;
        const double ms = mat.ms, ms_inv = std::fabs(ms) > 0.0 ? 1.0 / ms : 0.0;
        const double geom_eps = 1e-14;
        llgmi_mg_config ci = c;
        ci.warm_start = 0;
        ci.has_tol_abs = 0;
        ci.has_tol_rel = 0;
        ci.v_cycles = h.delta_v_cycles;
        ci.v_cycles_max = std::max<int64_t>(h.delta_v_cycles, 1);
        MGSolver imp(g, ci, stream, prof);
        ensure_fft();
        ensure_io();
        alloc(d_bd, 3 * n);
        T *tm = nullptr, *tb = nullptr;
        LLGMI_HIP(hipMalloc(&tm, sizeof(T) * 3 * n));
        std::vector<double> mh((size_t)(3 * n)), bf((size_t)(3 * n)), bm((size_t)(3 * n));
        std::vector<double> xx((size_t)nst), xy((size_t)nst), yy((size_t)nst), zz((size_t)nst);
        std::vector<double> xy_x((size_t)nst), xy_y((size_t)nst);
        try {
            LLGMI_HIP(hipMalloc(&tb, sizeof(T) * 3 * n));
            for (int axis = 0; axis < 3; ++axis) {
                for (int64_t q = 0; q < n; ++q) {
                    mh[(size_t)(3 * q)] = geom_eps;
                    mh[(size_t)(3 * q + 1)] = 0.0;
                    mh[(size_t)(3 * q + 2)] = 0.0;
                }
                const size_t ctr = (size_t)(3 * (cy * nx + cx));
                if (axis == 0) mh[ctr] = 1.0 + geom_eps;
                else if (axis == 1) mh[ctr + 1] = 1.0;
                else mh[ctr + 2] = 1.0;
                upload(mh.data(), tm);
                fft.field(tm, T(ms), d_bd, stream, prof);
                download(d_bd, bf.data());
                imp.template demag_field<T>(tm, ms, tb, h.sigma_cells, false);
                download(tb, bm.data());
                for (int64_t b = -r; b <= r; ++b)
                    for (int64_t a = -r; a <= r; ++a) {
                        const size_t id = (size_t)(3 * ((cy + b) * nx + (cx + a)));
                        const size_t k = (size_t)((b + r) * st + (a + r));
                        if (axis == 0) {
                            xx[k] = (bf[id] - bm[id]) * ms_inv;
                            xy_x[k] = (bf[id + 1] - bm[id + 1]) * ms_inv;
                        } else if (axis == 1) {
                            yy[k] = (bf[id + 1] - bm[id + 1]) * ms_inv;
                            xy_y[k] = (bf[id] - bm[id]) * ms_inv;
                        } else {
                            zz[k] = (bf[id + 2] - bm[id + 2]) * ms_inv;
                        }
                    }
            }
        } catch (...) {
            (void)hipFree(tm);
            if (tb) (void)hipFree(tb);
            throw;
        }
        (void)hipFree(tm);
        (void)hipFree(tb);
        for (int64_t k = 0; k < nst; ++k) xy[(size_t)k] = 0.5 * (xy_x[(size_t)k] + xy_y[(size_t)k]);
        // DeltaKernel2D::symmetrize (:664-694), in place in the reference's order
        auto ki = [&](int64_t a, int64_t b) { return (size_t)((b + r) * st + (a + r)); };
        for (int64_t b = -r; b <= r; ++b)
            for (int64_t a = -r; a <= r; ++a) {
                const size_t k = ki(a, b), kr = ki(-a, -b);
                xx[k] = 0.5 * (xx[k] + xx[kr]);
                yy[k] = 0.5 * (yy[k] + yy[kr]);
                zz[k] = 0.5 * (zz[k] + zz[kr]);
                xy[k] = 0.5 * (xy[k] + xy[kr]);
            }
        if (std::fabs(g.dx - g.dy) <= 1e-12 * std::max({std::fabs(g.dx), std::fabs(g.dy), 1.0})) {
            for (int64_t b = -r; b <= r; ++b)
                for (int64_t a = -r; a <= r; ++a) {
                    const size_t k = ki(a, b), ks = ki(b, a);
                    const double nxx = 0.5 * (xx[k] + yy[ks]), nyy = 0.5 * (yy[k] + xx[ks]),
                                 nxy = 0.5 * (xy[k] + xy[ks]);
                    xx[k] = nxx;
                    yy[k] = nyy;
                    xy[k] = nxy;
                }
        }
        // uniform-bias fix on the diagonal stencils (:3562-3571)
        const size_t center = (size_t)(r * st + r);
        double sxx = 0.0, syy = 0.0, szz = 0.0;
        for (double v : xx) sxx += v;
        for (double v : yy) syy += v;
        for (double v : zz) szz += v;
        xx[center] -= sxx;
        yy[center] -= syy;
        zz[center] -= szz;
        dk_host.assign((size_t)(4 * nst), 0.0);
        std::copy(xx.begin(), xx.end(), dk_host.begin());
        std::copy(xy.begin(), xy.end(), dk_host.begin() + nst);
        std::copy(yy.begin(), yy.end(), dk_host.begin() + 2 * nst);
        std::copy(zz.begin(), zz.end(), dk_host.begin() + 3 * nst);
        upload_dk();
    }

    LocalArgs<T> local_args(const llgmi_params &p, const llgmi_material &mat, int mask, bool torque,
                            bool relax = false) const {
        LocalArgs<T> a{};
        a.nx = (int)g.nx;
        a.ny = (int)g.ny;
        for (int q = 0; q < 3; ++q) a.bext[q] = T(p.b_ext[q]);
        a.exch_on = (mat.a_ex != 0.0 && mat.ms != 0.0) ? 1 : 0;
        a.exch_coeff = T(mat.ms != 0.0 ? 2.0 * mat.a_ex / mat.ms : 0.0);
        a.inv_dx2 = T(1.0 / (g.dx * g.dx));
        a.inv_dy2 = T(1.0 / (g.dy * g.dy));
        bool dnz = mat.has_dmi && mat.dmi != 0.0;
        a.has_eta = (dnz && mat.a_ex != 0.0) ? 1 : 0;
        a.eta = T(a.has_eta ? mat.dmi / (2.0 * mat.a_ex) : 0.0);
        a.dx = T(g.dx);
        a.dy = T(g.dy);
        a.anis_on = (mat.k_u != 0.0 && mat.ms != 0.0) ? 1 : 0;
        a.anis_coeff = T(mat.ms != 0.0 ? 2.0 * mat.k_u / mat.ms : 0.0);
        for (int q = 0; q < 3; ++q) a.u[q] = T(mat.easy_axis[q]);
        bool inc_dmi = mask == LLGMI_MASK_EXCH_ANIS_DMI || mask == LLGMI_MASK_FULL;
        a.dmi_on = (inc_dmi && dnz && mat.ms != 0.0 && mat.a_ex != 0.0) ? 1 : 0;
        a.dmi_pref = T(mat.ms != 0.0 ? 2.0 * mat.dmi / mat.ms : 0.0);
        a.inv_2dx = T(1.0 / (2.0 * g.dx));
        a.inv_2dy = T(1.0 / (2.0 * g.dy));
        a.torque = torque ? (relax ? 2 : 1) : 0;
        a.rhs_pref = relax ? T(-p.gamma * p.alpha / (1.0 + p.alpha * p.alpha))
                           : T(-p.gamma / (1.0 + p.alpha * p.alpha));
        a.alpha = T(p.alpha);
        return a;
    }

    // build_h_eff_masked[_geom] on device buffers; optional torque output.
    // relax: kout = llg_rhs_relax instead of llg_rhs.  b null: the field is
    // not stored (only the torque is wanted).  nxt (with kout): the next RK
    // stage state normalize(nxt_m + sum_q c_q k_q) is formed from the torque
    // in the same kernel and written to nxt_out, whose last k_q is kout.
    void field(const T *m, T *b, const llgmi_params &p, const llgmi_material &mat, int mask,
               const uint8_t *geom, T *kout, bool relax = false, const StageArgs<T> *nxt = nullptr,
               const T *nxt_m = nullptr, T *nxt_out = nullptr) {
        const bool demag = mask == LLGMI_MASK_FULL && mat.demag;
        LocalArgs<T> a = local_args(p, mat, mask, kout != nullptr, relax);
        if (nxt) {
            if (!kout || relax || nxt->nk < 1 || nxt->nk > 5 || nxt->k[nxt->nk - 1] != kout)
                throw Error(LLGMI_ERR_INVALID, "fused RK stage: the last slope must be this evaluation's torque");
            a.st_out = nxt_out;
            a.st_m = nxt_m;
            a.st_nk = nxt->nk;
            a.st_dt = nxt->dt;
            for (int q = 0; q < nxt->nk; ++q) {
                if (q < nxt->nk - 1) a.st_k[q] = nxt->k[q];
                a.st_c[q] = nxt->c[q];
                a.st_a[q] = nxt->a[q];
            }
        }
        if (demag && resolved_demag_method(mat) == LLGMI_DEMAG_FFT) {
            // FFT demag: inverse row pass and local field share one kernel
            ensure_fft();
            alloc(d_bd, 3 * n);
            fft.field_local(m, T(mat.ms), a, geom, b, kout, d_bd, stream, prof);
            return;
        }
        if (demag) demag_into_bd(m, mat, resolved_demag_method(mat), 0, 0);
        ProfScope ps(prof, LLGMI_K_LOCAL_FIELD, stream);
        launch_local_field<T>(a, m, demag ? d_bd : nullptr, geom, b, kout, stream);
    }

    // field() with every buffer beyond m and b taken from `scr` (scr_bytes
    // long): the FFT spectrum and, when the inverse row pass runs unfused,
    // B_demag.  Nothing is allocated or built: the FFT operator must exist and
    // be tuned, and the multigrid demag is not offered.  Not profiled.
    void field_noalloc(const T *m, T *b, const llgmi_params &p, const llgmi_material &mat, int mask,
                       const uint8_t *geom, unsigned char *scr, size_t scr_bytes) {
        const bool demag = mask == LLGMI_MASK_FULL && mat.demag;
        LocalArgs<T> a = local_args(p, mat, mask, false, false);
        if (!demag) {
            launch_local_field<T>(a, m, nullptr, geom, b, nullptr, stream);
            return;
        }
        if (resolved_demag_method(mat) != LLGMI_DEMAG_FFT)
            throw Error(LLGMI_ERR_UNSUPPORTED, "the allocation-free field evaluation offers the FFT demag only");
        if (!fft.built || fft.rq_fused == 0)
            throw Error(LLGMI_ERR_INVALID, "the context's FFT demag operator is not set up: call "
                                           "llgmi_dev_build_h_eff (Full mask, FFT demag) once first");
        const size_t spec_bytes = sizeof(cpx<T>) * 3 * (size_t)(fft.px / 2) * (size_t)fft.ny;
        const size_t bd_bytes = fft.rq_fused < 0 ? sizeof(T) * 3 * (size_t)n : 0;
        const size_t need = ((spec_bytes + 255) & ~size_t(255)) + bd_bytes;
        if (need > scr_bytes)
            throw Error(LLGMI_ERR_UNSUPPORTED, "the allocation-free field evaluation needs " + std::to_string(need) +
                                                   " bytes of scratch for this grid; the maximum is " +
                                                   std::to_string(scr_bytes));
        cpx<T> *own = fft.spec;
        fft.spec = reinterpret_cast<cpx<T> *>(scr);
        T *bd = bd_bytes ? reinterpret_cast<T *>(scr + ((spec_bytes + 255) & ~size_t(255))) : nullptr;
        try {
            fft.field_local(m, T(mat.ms), a, geom, b, nullptr, bd, stream, nullptr);
        } catch (...) {
            fft.spec = own;
            throw;
        }
        fft.spec = own;
    }

    // The staging buffers must exist before a destination is named: callers
    // pass d_m etc. by value, so ensure_io() runs before the call (a null dst
    // is rejected here rather than written through on the device).
    void upload(const double *h_aos, T *dst) {
        if (!dst || !d_aos) throw Error(LLGMI_ERR_INVALID, "internal: upload before allocation");
        LLGMI_HIP(hipMemcpyAsync(d_aos, h_aos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, stream));
        launch_aos_to_soa<T>(d_aos, dst, n, stream);
    }
    void download(const T *src, double *h_aos) {
        if (!src || !d_aos) throw Error(LLGMI_ERR_INVALID, "internal: download before allocation");
        launch_soa_to_aos<T>(src, d_aos, n, stream);
        LLGMI_HIP(hipMemcpyAsync(h_aos, d_aos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, stream));
        LLGMI_HIP(hipStreamSynchronize(stream));
    }
    const uint8_t *upload_geom(const uint8_t *geom) {
        if (!geom) return nullptr;
        alloc(d_geom, n);
        LLGMI_HIP(hipMemcpyAsync(d_geom, geom, (size_t)n, hipMemcpyHostToDevice, stream));
        return d_geom;
    }

    void host_build(const double *m, double *b, const llgmi_params &p, const llgmi_material &mat,
                    int mask, const uint8_t *geom) {
        ensure_io();
        upload(m, d_m);
        const uint8_t *dg = upload_geom(geom);
        field(d_m, d_b, p, mat, mask, dg, nullptr);
        download(d_b, b);
    }

    void host_add_demag(const double *m, double *b, const llgmi_material &mat, uint64_t pbc_x,
                        uint64_t pbc_y, int method = -1) {
        if (!mat.demag) return;
        ensure_io();
        upload(m, d_m);
        demag_into_bd(d_m, mat, method >= 0 ? method : resolved_demag_method(mat), pbc_x, pbc_y);
        h_tmp.resize((size_t)(3 * n));
        download(d_bd, h_tmp.data());
        for (int64_t i = 0; i < 3 * n; ++i) b[i] += h_tmp[(size_t)i];
    }

    void ensure_rk() {
        for (int s = 1; s <= 7; ++s) alloc(d_k[s], 3 * n);
        alloc(d_stage, 3 * n);
        alloc(d_stage2, 3 * n);
        alloc(d_m7, 3 * n);
        alloc(d_b, 3 * n);
        alloc(d_partial, kErrBlocks);
        alloc(d_err, 1);
        alloc(d_dt, 1);
        if (!h_err) LLGMI_HIP(hipHostMalloc((void **)&h_err, sizeof(double), hipHostMallocDefault));
        if (!h_dt) LLGMI_HIP(hipHostMalloc((void **)&h_dt, sizeof(double), hipHostMallocDefault));
    }

    // step_llg_rk45_recompute_field_masked_adaptive (llg.rs:795-1026) on device state m.
    void rk45(T *m, llgmi_params &p, const llgmi_material &mat, int mask, double max_err,
              double headroom, double dt_min, double dt_max, double &eps_out, int &acc_out,
              double &dtu_out) {
        rk45_issue(m, p, mat, mask);
        LLGMI_HIP(hipStreamSynchronize(stream));
        rk45_finish(m, p, mask, max_err, headroom, dt_min, dt_max, eps_out, acc_out, dtu_out);
    }

    // First half of an attempt: the stages and the error norm, ending with an
    // asynchronous copy of max|err| into pinned memory.  Nothing waits, so the
    // attempts of many contexts (one stream each) run side by side; p is read,
    // not changed.
    void rk45_issue(T *m, const llgmi_params &p, const llgmi_material &mat, int mask) {
        ensure_rk();
        bool fsal_ok = fsal_valid && last_mask == mask && last_bext[0] == p.b_ext[0] &&
                       last_bext[1] == p.b_ext[1] && last_bext[2] == p.b_ext[2];
        if (fsal_ok) {
            std::swap(d_k[1], d_k[7]);
            k_swapped = !k_swapped;
        }
        fsal_valid = false;   // set again by rk45_finish on accept; an attempt cut short leaves it off
        const bool demag = mask == LLGMI_MASK_FULL && mat.demag;
        const bool capturable = (!demag || resolved_demag_method(mat) == LLGMI_DEMAG_FFT) && prof->mask == 0 &&
                                rk_graphs_enabled();
        if (!capturable) {
            rk45_launches(m, p, mat, mask, fsal_ok, nullptr);
        } else {
            // one eager attempt per launch configuration (lazy set-up happens
            // there), then capture on the next and replay after that
            RkGraph &G = rk_graph[fsal_ok][k_swapped];
            const std::string key = rk_graph_key(m, p, mat, mask);
            *h_dt = p.dt;   // the graph's first node copies it (pinned memory, read when the graph runs)
            if (G.exec && G.key == key) {
                LLGMI_HIP(hipGraphLaunch(G.exec, stream));
            } else if (G.seen == key) {
                if (G.exec) (void)hipGraphExecDestroy(G.exec);
                G.exec = nullptr;
                G.key.clear();
                hipGraph_t graph = nullptr;
                LLGMI_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                try {
                    rk45_launches(m, p, mat, mask, fsal_ok, d_dt);
                } catch (...) {
                    (void)hipStreamEndCapture(stream, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                LLGMI_HIP(hipStreamEndCapture(stream, &graph));
                const hipError_t e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                LLGMI_HIP(e);
                G.key = key;
                LLGMI_HIP(hipGraphLaunch(G.exec, stream));
            } else {
                G.seen = key;
                rk45_launches(m, p, mat, mask, fsal_ok, d_dt);
            }
        }
        issued_mask = mask;
    }

    // The launches of one attempt: stages 2..7 with their field evaluations
    // (stage 1 too without FSAL), the FP64 error norm and the copy of
    // max|err| into pinned memory.  With dt_dev the step size is first copied
    // there from pinned h_dt and the stage coefficients are formed on the
    // device from it (for graph replay: the copy node reads h_dt when it runs).
    void rk45_launches(T *m, const llgmi_params &p, const llgmi_material &mat, int mask, bool fsal_ok,
                       double *dt_dev) {
        if (dt_dev) LLGMI_HIP(hipMemcpyAsync(dt_dev, h_dt, sizeof(double), hipMemcpyHostToDevice, stream));
        if (!fsal_ok) field(m, d_b, p, mat, mask, nullptr, d_k[1]);
        rk45_stages(m, p, mat, mask, dt_dev);
        {
            ProfScope ps(prof, LLGMI_K_RK45_ERR, stream);
            launch_rk_err<T>(d_k[1], d_k[3], d_k[4], d_k[5], d_k[6], d_k[7], n,
                             sizeof(T) == sizeof(float), d_partial, d_err, stream);
        }
        LLGMI_HIP(hipMemcpyAsync(h_err, d_err, sizeof(double), hipMemcpyDeviceToHost, stream));
    }

    // Stages 2..7 with their field evaluations (k1 in place); with dt_dev the
    // stage kernels take the step size from device memory.  The state of
    // stage s + 1 is formed in the epilogue of stage s's field evaluation
    // from its torque (stage 2's by the stage kernel; LLGMI_RK45_FUSE=0 runs
    // every stage combination as its own kernel, with identical results).
    // Stage states alternate between two buffers, so an epilogue never writes
    // the state its own field evaluation reads.  The field b itself is not
    // stored.  acc_flag (device-controlled runs): when the previous attempt
    // was accepted, the stage-2 kernel first takes m = m7 and k1 = k7.
    void rk45_stages(T *m, const llgmi_params &p, const llgmi_material &mat, int mask, double *dt_dev,
                     const int *acc_flag = nullptr) {
        static const double A[7][5] = {
            {0}, {0},
            {1.0 / 5.0},
            {3.0 / 40.0, 9.0 / 40.0},
            {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0},
            {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0},
            {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0},
        };
        static const double Bw[5] = {35.0 / 384.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0};
        static const int Bk[5] = {1, 3, 4, 5, 6};
        const double dt0 = p.dt;
        auto args = [&](int s) {
            StageArgs<T> sa{};
            sa.dt = dt_dev;
            if (s < 7) {
                sa.nk = s - 1;
                for (int j = 1; j < s; ++j) {
                    sa.k[j - 1] = d_k[j];
                    sa.a[j - 1] = A[s][j - 1];
                    sa.c[j - 1] = T(dt0 * A[s][j - 1]);
                }
            } else {
                sa.nk = 5;
                for (int q = 0; q < 5; ++q) {
                    sa.k[q] = d_k[Bk[q]];
                    sa.a[q] = Bw[q];
                    sa.c[q] = T(dt0 * Bw[q]);
                }
            }
            return sa;
        };
        auto state = [&](int s) { return s == 7 ? d_m7 : ((s & 1) ? d_stage2 : d_stage); };
        auto stage_kernel = [&](int s) {
            const StageArgs<T> sa = args(s);
            ProfScope ps(prof, LLGMI_K_RK45_STAGE, stream);
            launch_rk_stage<T>(m, sa, state(s), n, stream);
        };
        const bool fuse = rk45_fuse_enabled();
        if (acc_flag) {
            const StageArgs<T> sa = args(2);
            ProfScope ps(prof, LLGMI_K_RK45_STAGE, stream);
            launch_rk_stage2_accept<T>(acc_flag, m, d_m7, d_k[1], d_k[7], sa, state(2), n, stream);
        } else {
            stage_kernel(2);
        }
        for (int s = 2; s <= 7; ++s) {
            if (s < 7 && fuse) {
                const StageArgs<T> nxt = args(s + 1);
                field(state(s), nullptr, p, mat, mask, nullptr, d_k[s], false, &nxt, m, state(s + 1));
            } else {
                field(state(s), nullptr, p, mat, mask, nullptr, d_k[s]);
                if (s < 7) stage_kernel(s + 1);
            }
        }
    }

    static bool rk45_fuse_enabled() {
        const char *e = std::getenv("LLGMI_RK45_FUSE");
        return !(e && e[0] == '0');
    }
    static bool rk_graphs_enabled() {
        const char *e = std::getenv("LLGMI_RK45_GRAPH");
        return !(e && e[0] == '0');
    }
    // everything a captured attempt reads from the host besides the step size
    std::string rk_graph_key(const T *m, const llgmi_params &p, const llgmi_material &mat, int mask) const {
        std::string k;
        auto put = [&](const void *v, size_t b) { k.append((const char *)v, b); };
        put(&m, sizeof m);
        put(&mask, sizeof mask);
        put(&p.gamma, sizeof p.gamma);
        put(&p.alpha, sizeof p.alpha);
        put(p.b_ext, sizeof p.b_ext);
        put(&mat, sizeof mat);
        put(d_k, sizeof d_k);
        put(&d_stage, sizeof d_stage);
        put(&d_stage2, sizeof d_stage2);
        const bool fuse = rk45_fuse_enabled();
        put(&fuse, sizeof fuse);
        put(&d_m7, sizeof d_m7);
        put(&d_b, sizeof d_b);
        const int method = resolved_demag_method(mat);
        put(&method, sizeof method);
        put(&fft_gen, sizeof fft_gen);
        return k;
    }

    // Second half, after the stream has completed rk45_issue's work: the
    // controller (llg.rs:999-1025), FP64 on the host, and the accepted state.
    void rk45_finish(T *m, llgmi_params &p, int mask, double max_err, double headroom, double dt_min,
                     double dt_max, double &eps_out, int &acc_out, double &dtu_out) {
        rk45_decide(m, p, mask, max_err, headroom, dt_min, dt_max, *h_err, true, eps_out, acc_out, dtu_out);
    }

    // The attempt's FSAL decision and scratch state without launching anything
    // (the batched stepper launches the stages of many films itself): the
    // state changes rk45_issue makes.  Returns whether k1 is reused.
    bool rk45_prepare(const llgmi_params &p, const llgmi_material &mat, int mask) {
        ensure_rk();
        const bool fsal_ok = fsal_valid && last_mask == mask && last_bext[0] == p.b_ext[0] &&
                             last_bext[1] == p.b_ext[1] && last_bext[2] == p.b_ext[2];
        if (fsal_ok) {
            std::swap(d_k[1], d_k[7]);
            k_swapped = !k_swapped;
        }
        fsal_valid = false;
        if (mask == LLGMI_MASK_FULL && mat.demag && resolved_demag_method(mat) == LLGMI_DEMAG_FFT) {
            ensure_fft();
            alloc(d_bd, 3 * n);
        }
        issued_mask = mask;
        return fsal_ok;
    }

    // The controller (llg.rs:999-1025) from the attempt's max |err|; with copy
    // the accepted state m7 is copied into m on the context stream (the
    // batched stepper copies the accepted films itself).
    void rk45_decide(T *m, llgmi_params &p, int mask, double max_err, double headroom, double dt_min,
                     double dt_max, double err_inf, bool copy, double &eps_out, int &acc_out, double &dtu_out) {
        if (issued_mask != mask || mask < 0) throw Error(LLGMI_ERR_INVALID, "internal: rk45_finish without rk45_issue");
        issued_mask = -1;
        const double dt0 = p.dt;
        const double eps = err_inf * dt0;
        double dt_next = eps > 0.0 ? dt0 * headroom * std::pow(max_err / eps, 0.2) : dt0 * 2.0;
        if (dt_next < dt_min) dt_next = dt_min;
        if (dt_next > dt_max) dt_next = dt_max;
        const bool accept = (eps <= max_err) || (dt0 <= dt_min * 1.0000000001);
        if (accept) {
            if (copy) LLGMI_HIP(hipMemcpyAsync(m, d_m7, sizeof(T) * 3 * n, hipMemcpyDeviceToDevice, stream));
            fsal_valid = true;
            last_mask = mask;
            for (int q = 0; q < 3; ++q) last_bext[q] = p.b_ext[q];
        } else {
            fsal_valid = false;
            last_mask = -1;
        }
        p.dt = dt_next;
        eps_out = eps;
        acc_out = accept ? 1 : 0;
        dtu_out = dt0;
    }

    // `attempts` consecutive attempts with the controller (llg.rs:999-1025,
    // FP64) on the device, replayed from one captured graph: per attempt the
    // stages and the error norm with the controller in its final reduction
    // (eps, decision, next step size).  After an accepted attempt m = m7 and
    // k1 = k7, applied by the next attempt's stage-2 kernel (k_rk_accept after
    // the last); after a rejected one m and k1 = f(m) stay, which is what
    // recomputing k1 gives.  The host waits once; p.dt returns the next step
    // size.
    void rk45_run(T *m, llgmi_params &p, const llgmi_material &mat, int mask, double max_err, double headroom,
                  double dt_min, double dt_max, int64_t attempts, double *eps, int *acc, double *dtu) {
        const bool demag = mask == LLGMI_MASK_FULL && mat.demag;
        if (demag && resolved_demag_method(mat) != LLGMI_DEMAG_FFT)
            throw Error(LLGMI_ERR_UNSUPPORTED, "device-controlled runs need the FFT demag (or none)");
        if (prof->mask != 0) throw Error(LLGMI_ERR_UNSUPPORTED, "device-controlled runs need kernel profiling off");
        ensure_rk();
        alloc(d_run_flag, 1);
        if (run_cap < attempts) {
            drop_rk_graphs();
            free_run_outputs();
            LLGMI_HIP(hipMalloc(&d_run_eps, sizeof(double) * attempts));
            LLGMI_HIP(hipMalloc(&d_run_dtu, sizeof(double) * attempts));
            LLGMI_HIP(hipMalloc(&d_run_acc, sizeof(int) * attempts));
            LLGMI_HIP(hipHostMalloc((void **)&h_run_eps, sizeof(double) * attempts, hipHostMallocDefault));
            LLGMI_HIP(hipHostMalloc((void **)&h_run_dtu, sizeof(double) * attempts, hipHostMallocDefault));
            LLGMI_HIP(hipHostMalloc((void **)&h_run_acc, sizeof(int) * attempts, hipHostMallocDefault));
            run_cap = attempts;
        }
        if (demag) {
            // lazy set-up and tuning happen outside the capture
            ensure_fft();
            alloc(d_bd, 3 * n);
            if (fft.rq_fused == 0) {
                field(m, d_b, p, mat, mask, nullptr, nullptr);
                LLGMI_HIP(hipStreamSynchronize(stream));
            }
        }
        const bool fsal_ok = fsal_valid && last_mask == mask && last_bext[0] == p.b_ext[0] &&
                             last_bext[1] == p.b_ext[1] && last_bext[2] == p.b_ext[2];
        if (fsal_ok) {
            std::swap(d_k[1], d_k[7]);
            k_swapped = !k_swapped;
        }
        fsal_valid = false;
        std::string key = rk_graph_key(m, p, mat, mask);
        auto put = [&](const void *v, size_t b) { key.append((const char *)v, b); };
        const double ctl[4] = {max_err, headroom, dt_min, dt_max};
        put(&attempts, sizeof attempts);
        put(&fsal_ok, sizeof fsal_ok);
        put(ctl, sizeof ctl);
        put(&d_run_eps, sizeof d_run_eps);
        *h_dt = p.dt;   // the graph's first node copies it
        if (!run_graph.exec || run_graph.key != key) {
            if (run_graph.exec) (void)hipGraphExecDestroy(run_graph.exec);
            run_graph = RkGraph{};
            hipGraph_t graph = nullptr;
            LLGMI_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                LLGMI_HIP(hipMemcpyAsync(d_dt, h_dt, sizeof(double), hipMemcpyHostToDevice, stream));
                if (!fsal_ok) field(m, d_b, p, mat, mask, nullptr, d_k[1]);
                for (int64_t a = 0; a < attempts; ++a) {
                    rk45_stages(m, p, mat, mask, d_dt, a > 0 ? d_run_flag : nullptr);
                    RkControl ctl;
                    ctl.dt = d_dt;
                    ctl.flag = d_run_flag;
                    ctl.max_err = max_err;
                    ctl.headroom = headroom;
                    ctl.dt_min = dt_min;
                    ctl.dt_max = dt_max;
                    ctl.a = (int)a;
                    ctl.o_eps = d_run_eps;
                    ctl.o_acc = d_run_acc;
                    ctl.o_dtu = d_run_dtu;
                    launch_rk_err<T>(d_k[1], d_k[3], d_k[4], d_k[5], d_k[6], d_k[7], n,
                                     sizeof(T) == sizeof(float), d_partial, d_err, stream, &ctl);
                }
                // the last attempt's accept (each earlier one is applied by the next stage 2)
                launch_rk_accept<T>(d_run_flag, m, d_m7, d_k[1], d_k[7], n, stream);
                LLGMI_HIP(hipMemcpyAsync(h_run_eps, d_run_eps, sizeof(double) * attempts, hipMemcpyDeviceToHost,
                                         stream));
                LLGMI_HIP(hipMemcpyAsync(h_run_dtu, d_run_dtu, sizeof(double) * attempts, hipMemcpyDeviceToHost,
                                         stream));
                LLGMI_HIP(hipMemcpyAsync(h_run_acc, d_run_acc, sizeof(int) * attempts, hipMemcpyDeviceToHost,
                                         stream));
                LLGMI_HIP(hipMemcpyAsync(h_dt, d_dt, sizeof(double), hipMemcpyDeviceToHost, stream));
            } catch (...) {
                (void)hipStreamEndCapture(stream, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            LLGMI_HIP(hipStreamEndCapture(stream, &graph));
            const hipError_t ge = hipGraphInstantiate(&run_graph.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            LLGMI_HIP(ge);
            run_graph.key = key;
        }
        LLGMI_HIP(hipGraphLaunch(run_graph.exec, stream));
        LLGMI_HIP(hipStreamSynchronize(stream));
        std::memcpy(eps, h_run_eps, sizeof(double) * attempts);
        std::memcpy(dtu, h_run_dtu, sizeof(double) * attempts);
        std::memcpy(acc, h_run_acc, sizeof(int) * attempts);
        p.dt = *h_dt;
        // k1 holds f(m) after the last attempt, accepted or not: the engine
        // keeps it in the k7 slot, as after an accepted single attempt
        std::swap(d_k[1], d_k[7]);
        k_swapped = !k_swapped;
        fsal_valid = true;
        last_mask = mask;
        for (int q = 0; q < 3; ++q) last_bext[q] = p.b_ext[q];
    }

    // ---- relaxation ------------------------------------------------------------
    void ensure_red() {
        alloc(d_red, (size_t)5 * kRelaxRedBlocks);
        alloc(d_redout, 8);
        if (!h_red) LLGMI_HIP(hipHostMalloc((void **)&h_red, 8 * sizeof(double), hipHostMallocDefault));
    }
    void read_red(int count) {
        LLGMI_HIP(hipMemcpyAsync(h_red, d_redout, sizeof(double) * count, hipMemcpyDeviceToHost, stream));
        LLGMI_HIP(hipStreamSynchronize(stream));
    }
    void ensure_rk23() {
        for (int s = 1; s <= 4; ++s) alloc(d_r23k[s], 3 * n);
        alloc(d_r23b4, 3 * n);
        alloc(d_r23m4, 3 * n);
        alloc(d_stage, 3 * n);
        alloc(d_b, 3 * n);
        ensure_red();
    }
    void rk23_invalidate() {
        r23_b4_valid = false;
        r23_has_last = false;
        r23_last_mask = -1;
    }

    // step_llg_rk23_recompute_field_masked_relax_adaptive (llg.rs:278-436) on device state m.
    void rk23(T *m, llgmi_params &p, const llgmi_material &mat, int mask, double max_err,
              double headroom, double dt_min, double dt_max, double &eps_out, int &acc_out,
              double &dtu_out) {
        const double A21 = 1.0 / 2.0, A32 = 3.0 / 4.0;
        const double B1 = 2.0 / 9.0, B2 = 1.0 / 3.0, B3 = 4.0 / 9.0;
        const double BH1 = 7.0 / 24.0, BH2 = 1.0 / 4.0, BH3 = 1.0 / 3.0, BH4 = 1.0 / 8.0;
        ensure_rk23();
        T **K = d_r23k;
        const double dt0 = p.dt;
        const bool fsal_ok = r23_b4_valid && r23_has_last && r23_last_mask == mask &&
                             r23_last_bext[0] == p.b_ext[0] && r23_last_bext[1] == p.b_ext[1] &&
                             r23_last_bext[2] == p.b_ext[2];
        // FSAL: k1 <- k4 of the accepted state (b1 is not read again)
        if (fsal_ok) std::swap(K[1], K[4]);
        else field(m, d_b, p, mat, mask, nullptr, K[1], true);
        auto stage = [&](std::initializer_list<std::pair<const T *, double>> terms, T *out) {
            StageArgs<T> sa{};
            sa.nk = 0;
            for (auto &t : terms) {
                sa.k[sa.nk] = t.first;
                sa.c[sa.nk] = T(dt0 * t.second);
                ++sa.nk;
            }
            ProfScope ps(prof, LLGMI_K_RELAX, stream);
            launch_rk_stage<T>(m, sa, out, n, stream);
        };
        stage({{K[1], A21}}, d_stage);
        field(d_stage, d_b, p, mat, mask, nullptr, K[2], true);
        stage({{K[2], A32}}, d_stage);
        field(d_stage, d_b, p, mat, mask, nullptr, K[3], true);
        stage({{K[1], B1}, {K[2], B2}, {K[3], B3}}, d_r23m4);
        field(d_r23m4, d_r23b4, p, mat, mask, nullptr, K[4], true);
        {
            Rk23ErrCoef c{{dt0 * B1, dt0 * B2, dt0 * B3}, {dt0 * BH1, dt0 * BH2, dt0 * BH3, dt0 * BH4}};
            ProfScope ps(prof, LLGMI_K_RELAX, stream);
            launch_rk23_err<T>(m, K[1], K[2], K[3], K[4], d_r23m4, c, n, d_red, d_redout, stream);
        }
        read_red(1);
        const double eps = h_red[0];
        // controller (llg.rs:422-435), FP64 on the host
        double dt_next = eps > 0.0 ? dt0 * headroom * std::pow(max_err / eps, 1.0 / 3.0) : dt0 * 2.0;
        if (dt_next < dt_min) dt_next = dt_min;
        if (dt_next > dt_max) dt_next = dt_max;
        const bool accept = (eps <= max_err) || (dt0 <= dt_min * 1.0000000001);
        if (accept) {
            LLGMI_HIP(hipMemcpyAsync(m, d_r23m4, sizeof(T) * 3 * n, hipMemcpyDeviceToDevice, stream));
            r23_b4_valid = true;
            r23_has_last = true;
            r23_last_mask = mask;
            for (int q = 0; q < 3; ++q) r23_last_bext[q] = p.b_ext[q];
        } else {
            rk23_invalidate();
        }
        p.dt = dt_next;
        eps_out = eps;
        acc_out = accept ? 1 : 0;
        dtu_out = dt0;
    }

    // compute_energy_geom (energy.rs:80-283); geom on the device or null.
    void energy(const T *m, const llgmi_material &mat, const double bext[3], const uint8_t *geom,
                double out[5]) {
        ensure_red();
        const T *bd = nullptr;
        if (mat.demag) {
            demag_into_bd(m, mat, resolved_demag_method(mat), 0, 0);
            bd = d_bd;
        }
        EnergyArgs a{};
        a.nx = (int)g.nx;
        a.ny = (int)g.ny;
        a.dx = g.dx;
        a.dy = g.dy;
        a.v = g.dx * g.dy * g.dz;
        a.aex = mat.a_ex;
        a.ku = mat.k_u;
        a.ms = mat.ms;
        for (int q = 0; q < 3; ++q) { a.u[q] = mat.easy_axis[q]; a.bext[q] = bext[q]; }
        a.dmi_on = (mat.has_dmi && mat.dmi != 0.0 && mat.a_ex != 0.0) ? 1 : 0;
        a.pref = mat.ms != 0.0 ? 2.0 * mat.dmi / mat.ms : 0.0;
        a.eta = a.dmi_on ? mat.dmi / (2.0 * mat.a_ex) : 0.0;
        {
            ProfScope ps(prof, LLGMI_K_RELAX, stream);
            launch_energy<T>(a, m, bd, geom, d_red, d_redout, stream);
        }
        read_red(5);
        for (int q = 0; q < 5; ++q) out[q] = h_red[q];
    }
    double energy_total(const T *m, const llgmi_material &mat, const double bext[3]) {
        double e[5];
        energy(m, mat, bext, nullptr, e);
        return e[0] + e[1] + e[2] + e[3] + e[4];   // EnergyBreakdown::total (energy.rs:30-34)
    }

    // torque_metric_from_field (relax.rs:144-190)
    double torque_metric(const T *m, const T *b, int metric) {
        ensure_red();
        {
            ProfScope ps(prof, LLGMI_K_RELAX, stream);
            launch_torque<T>(m, b, n, d_red, d_redout, stream);
        }
        read_red(3);
        const double nf = std::max((double)n, 1.0);
        switch (metric) {
            case LLGMI_TORQUE_MAX: return h_red[0];
            case LLGMI_TORQUE_MEAN: return h_red[1] / nf;
            case LLGMI_TORQUE_RMS: return std::sqrt(h_red[2] / nf);
            default: return h_red[2] / nf;
        }
    }

    // relax_with_report (relax.rs:334-774) on device state m.
    void relax(T *m, llgmi_params &p, const llgmi_material &mat, int mask, llgmi_relax_settings &st,
               llgmi_relax_report &rep);
};

// relax_with_report (relax.rs:334-774).  The controller is host logic over
// device steps; each step and each torque / energy check synchronises once to
// read its scalar.
template <typename T>
void Engine<T>::relax(T *m, llgmi_params &p, const llgmi_material &mat, int mask,
                      llgmi_relax_settings &st, llgmi_relax_report &rep) {
    const bool phase1 = st.phase1_enabled != 0, phase2 = st.phase2_enabled != 0;
    const int64_t energy_stride = std::max<int64_t>(st.energy_stride, 1);
    const int metric = st.torque_metric;
    const bool has_thr = st.has_torque_threshold != 0;
    const int64_t check_stride = std::max<int64_t>(st.torque_check_stride, 1);
    const double plateau_rel = st.torque_plateau_rel, plateau_abs = st.torque_plateau_abs;
    const int64_t plateau_checks = st.torque_plateau_checks;
    const int64_t plateau_min_checks = std::max<int64_t>(st.torque_plateau_min_checks, 1);
    const int64_t max_acc = st.max_accepted_steps;
    const bool has_gate = st.has_final_torque_gate != 0, has_gate_max = st.has_final_torque_gate_max != 0;
    const double headroom = st.headroom, dt_min = st.dt_min, dt_max = st.dt_max;
    double max_err = st.max_err;

    p.dt = std::min(std::max(p.dt, dt_min), dt_max);
    rk23_invalidate();

    int64_t accepted = 0, rejected = 0, checks = 0, rebuilds = 0;
    int last_stage = LLGMI_STAGE_NONE;
    bool has_ft = false, has_ftmax = false;
    double ft = 0.0, ftmax = 0.0;

    auto finish = [&](int stop, int stage) {
        st.max_err = max_err;
        rep.accepted_steps = accepted;
        rep.rejected_steps = rejected;
        rep.torque_checks = checks;
        rep.torque_field_rebuilds = rebuilds;
        rep.stop_reason = stop;
        rep.last_stage_stop = stage;
        rep.has_final_torque = has_ft ? 1 : 0;
        rep.final_torque = has_ft ? ft : 0.0;
        rep.has_final_torque_max = has_ftmax ? 1 : 0;
        rep.final_torque_max = has_ftmax ? ftmax : 0.0;
        rep.final_max_err = max_err;
        rep.final_dt = p.dt;
    };
    // torque_now (relax.rs:206-230)
    auto torque_now = [&](int which) {
        ++checks;
        if (r23_b4_valid) return torque_metric(m, d_r23b4, which);
        ++rebuilds;
        alloc(d_b, 3 * n);
        field(m, d_b, p, mat, mask, nullptr, nullptr);
        return torque_metric(m, d_b, which);
    };
    // advance_accepted (relax.rs:247-277)
    auto advance = [&](int64_t n_accept) {
        const int64_t target = accepted + std::max<int64_t>(n_accept, 1);
        while (accepted < target) {
            if (accepted >= max_acc) return false;
            double eps, dtu;
            int ok;
            rk23(m, p, mat, mask, max_err, headroom, dt_min, dt_max, eps, ok, dtu);
            if (ok) ++accepted;
            else ++rejected;
        }
        return true;
    };

    if (phase1) {
        double e0 = energy_total(m, mat, p.b_ext);
        for (;;) {
            if (!advance(energy_stride)) return finish(LLGMI_STOP_MAX_ACCEPTED_STEPS, last_stage);
            const double e1 = energy_total(m, mat, p.b_ext);
            const double tol = st.rel_energy_tol * std::max(std::fabs(e0), 1e-30);
            if (e1 < e0 - tol) {
                e0 = e1;
                continue;
            }
            break;
        }
    }
    if (!phase2) return finish(LLGMI_STOP_PHASE2_DISABLED, last_stage);

    const bool plateau_enabled = !has_thr && plateau_checks > 0;
    const int64_t need_fails = std::max<int64_t>(plateau_checks, 1);
    // MuMax's float32 sum-of-squares plateau (relax.rs:472-488)
    const bool mumax_f32 = plateau_enabled && metric == LLGMI_TORQUE_MEANSQ && plateau_rel == 0.0 &&
                           plateau_abs == 0.0 && plateau_checks == 1 && plateau_min_checks == 1;
    const double n_f = (double)n;
    for (;;) {
        double t_prev = torque_now(metric);
        float t_prev_f32 = mumax_f32 ? (float)(t_prev * n_f) : (float)t_prev;
        int64_t fails = 0, comparisons = 0;
        for (;;) {
            if (has_thr && t_prev <= st.torque_threshold) {
                last_stage = LLGMI_STAGE_BELOW_THRESHOLD;
                break;
            }
            if (plateau_enabled && fails >= need_fails) {
                last_stage = LLGMI_STAGE_PLATEAU;
                break;
            }
            if (accepted >= max_acc || !advance(check_stride)) {
                has_ft = true;
                ft = t_prev;
                return finish(LLGMI_STOP_MAX_ACCEPTED_STEPS, LLGMI_STAGE_MAX_ACCEPTED_STEPS);
            }
            const double t_new = torque_now(metric);
            if (plateau_enabled) {
                ++comparisons;
                if (comparisons < plateau_min_checks) {
                    fails = 0;
                } else if (mumax_f32) {
                    const float t_new_f32 = (float)(t_new * n_f);
                    const bool improved = t_new_f32 < t_prev_f32;
                    t_prev_f32 = t_new_f32;
                    fails = improved ? 0 : fails + 1;
                } else {
                    const double need = std::max(plateau_rel * std::max(std::fabs(t_prev), 1e-30), plateau_abs);
                    fails = (t_prev - t_new) > need ? 0 : fails + 1;
                }
            }
            t_prev = t_new;
        }
        has_ft = true;
        ft = t_prev;
        // tighten (MaxErr /= sqrt 2)
        if (max_err <= st.tighten_floor) break;
        max_err *= st.tighten_factor;
        if (max_err < st.tighten_floor) max_err = st.tighten_floor;
    }
    has_ftmax = true;
    ftmax = torque_now(LLGMI_TORQUE_MAX);

    if ((has_gate || has_gate_max) && st.gate_max_extra_accepted_steps > 0) {
        const int64_t start = accepted;
        double last_t = has_ft ? ft : INFINITY;
        int64_t gate_fails = 0;
        for (;;) {
            // gates_ok (relax.rs:279-332)
            const double t_metric = torque_now(metric);
            const bool need_max = has_gate_max || metric == LLGMI_TORQUE_MAX;
            double t_max = NAN;
            if (need_max) t_max = metric == LLGMI_TORQUE_MAX ? t_metric : torque_now(LLGMI_TORQUE_MAX);
            const bool ok = (!has_gate || t_metric <= st.final_torque_gate) &&
                            (!has_gate_max || t_max <= st.final_torque_gate_max);
            has_ft = true;
            ft = t_metric;
            if (std::isfinite(t_max)) { has_ftmax = true; ftmax = t_max; }
            if (ok) break;
            const int64_t used = accepted > start ? accepted - start : 0;
            if (used >= st.gate_max_extra_accepted_steps)
                return finish(LLGMI_STOP_TORQUE_GATE_NOT_SATISFIED, last_stage);
            if (st.gate_plateau_fails > 0) {
                const double need = std::max(plateau_rel * std::max(std::fabs(last_t), 1e-30), plateau_abs);
                gate_fails = (last_t - t_metric) > need ? 0 : gate_fails + 1;
                if (gate_fails >= st.gate_plateau_fails)
                    return finish(LLGMI_STOP_TORQUE_GATE_NOT_SATISFIED, last_stage);
                last_t = t_metric;
            }
            if (!advance(check_stride))
                return finish(LLGMI_STOP_MAX_ACCEPTED_STEPS, LLGMI_STAGE_MAX_ACCEPTED_STEPS);
        }
    }
    finish(LLGMI_STOP_TIGHTEN_FLOOR_REACHED, last_stage);
}

// Device scratch of llgmi_dev_build_h_eff_noalloc: one array per GPU, held
// until the process ends.  Host code never names the array itself (its host
// symbol is not a device address): each device's address comes from
// hipGetSymbolAddress, taken once per device when the first context on it is
// created.  Every call on a GPU shares that GPU's array; calls that overlap on
// one GPU race.
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

// with `device` current
void ensure_scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(LLGMI_ERR_UNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *ptr = nullptr;
    LLGMI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(LLGMI_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[device] = ptr;
}

void *scratch_address(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

struct llgmi_ctx {
    llgmi_grid grid{};
    int precision = LLGMI_F64;
    int device = 0;
    hipStream_t stream = nullptr;
    Profiler prof;
    std::unique_ptr<Engine<float>> ef;
    std::unique_ptr<Engine<double>> ed;
};

// ---------------------------------------------------------------------------
// argument checks
// ---------------------------------------------------------------------------
namespace {

std::string check_grid(const llgmi_grid *g) {
    if (!g) return "grid is NULL";
    if (g->nx <= 0 || g->ny <= 0) return "grid nx and ny must be positive";
    if (g->nx > (1 << 20) || g->ny > (1 << 20) || g->nx * g->ny > (int64_t(1) << 31))
        return "grid too large";
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0) || !std::isfinite(g->dx) ||
        !std::isfinite(g->dy) || !std::isfinite(g->dz))
        return "grid spacings must be positive and finite";
    return "";
}

std::string check_mat(const llgmi_material *m) {
    if (!m) return "material is NULL";
    if (m->demag_method < 0 || m->demag_method > 2) return "unknown demag_method";
    return "";
}

std::string check_mask(int mask) {
    if (mask < 0 || mask > 2) return "unknown field mask";
    return "";
}

// Device pointer check: the range [p, p+bytes) must be device memory owned by
// one allocation, so a host pointer or a short buffer is rejected instead of
// faulting the GPU.
std::string check_dev(const void *p, size_t bytes, const char *what) {
    if (!p) return std::string(what) + " is NULL";
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return std::string(what) + " is not a HIP device pointer";
    }
    if (at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged)
        return std::string(what) + " is not device memory";
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return std::string(what) + ": cannot query its allocation";
    }
    const char *b = (const char *)base, *q = (const char *)p;
    if (q < b || q + bytes > b + size) return std::string(what) + " is smaller than the grid needs";
    return "";
}

}  // namespace

std::string llgmi::check_device_range(const void *p, size_t bytes, const char *what) {
    return check_dev(p, bytes, what);
}

namespace {

template <typename F>
int guarded(llgmi_ctx *ctx, F &&f) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    try {
        LLGMI_HIP(hipSetDevice(ctx->device));
        f();
        return LLGMI_OK;
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(LLGMI_ERR_ALLOC, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(LLGMI_ERR_INVALID, e.what());
    }
}

template <typename F>
void with_engine(llgmi_ctx *ctx, F &&f) {
    if (ctx->precision == LLGMI_F32) f(*ctx->ef);
    else f(*ctx->ed);
}

size_t elem_size(const llgmi_ctx *ctx) { return ctx->precision == LLGMI_F32 ? 4 : 8; }

// ---------------------------------------------------------------------------
// Batched RK45 attempts of many films (llgmi_dev_rk45_step_batch): every
// kernel launch covers all films of one kernel class (films whose FFT passes
// have the same lengths and radix choices), through tables of (film, block)
// entries; the stage and error kernels cover all films at once.  One attempt
// of every film is one replay of a captured graph, then one host
// synchronisation for all controllers.  Each film runs exactly the kernels
// (bodies, radix choices, fused or unfused inverse pass) its own context runs
// for llgmi_dev_rk45_step, so its results are the single-film results.
// ---------------------------------------------------------------------------
template <typename T>
struct BatchPlan {
    struct Launch {
        int kind;   // 0..3: launch_batch_fft kinds; 4: local field (demag = with B_demag)
        int lg, rq;
        bool half;
        int demag;
        int off, n;   // entries in the table
    };
    int dev = -1;
    int nfilms = 0;
    hipStream_t stream = nullptr;
    BFilm<T> *d_films = nullptr, *h_films = nullptr;   // h_films pinned
    BEntry *d_tab = nullptr;
    std::vector<Launch> field;   // one field evaluation, in launch order
    int elem_off = 0, elem_n = 0;
    double *d_partial = nullptr, *d_err = nullptr, *h_err = nullptr;   // h_err pinned
    hipGraphExec_t exec = nullptr;
    // runs of several attempts with the controller on the device, one graph
    // per (attempts, controller settings); per-attempt outputs in device and
    // pinned host memory
    struct Run {
        hipGraphExec_t exec = nullptr;
        double *d_eps = nullptr, *d_dtu = nullptr, *h_eps = nullptr, *h_dtu = nullptr;
        int *d_acc = nullptr, *h_acc = nullptr;
    };
    std::map<std::tuple<int64_t, double, double, double, double>, Run> runs;

    ~BatchPlan() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &kv : runs) {
            Run &r = kv.second;
            if (r.exec) (void)hipGraphExecDestroy(r.exec);
            void *dv[] = {r.d_eps, r.d_dtu, r.d_acc};
            for (void *p : dv)
                if (p) (void)hipFree(p);
            void *hv[] = {r.h_eps, r.h_dtu, r.h_acc};
            for (void *p : hv)
                if (p) (void)hipHostFree(p);
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (d_films) (void)hipFree(d_films);
        if (d_tab) (void)hipFree(d_tab);
        if (d_partial) (void)hipFree(d_partial);
        if (d_err) (void)hipFree(d_err);
        if (h_films) (void)hipHostFree(h_films);
        if (h_err) (void)hipHostFree(h_err);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void launch_field(int which, int ks, int stage1) {
        for (const Launch &l : field) {
            if (l.kind < 4)
                launch_batch_fft<T>(l.kind, l.lg, l.rq, l.half, d_films, d_tab + l.off, l.n, which, ks, stage1, stream);
            else
                launch_batch_local<T>(d_films, d_tab + l.off, l.n, l.demag, which, ks, stage1, stream);
        }
    }
    // the attempt: H2D of the film records, stage 1 (films without FSAL),
    // stages 2..7 with their field evaluations, the error norms, D2H of them
    void launch_attempt() {
        LLGMI_HIP(hipMemcpyAsync(d_films, h_films, sizeof(BFilm<T>) * nfilms, hipMemcpyHostToDevice, stream));
        launch_field(0, 1, 1);
        for (int st = 2; st <= 7; ++st) {
            launch_batch_stage<T>(d_films, d_tab + elem_off, elem_n, st, stream);
            launch_field(st == 7 ? 2 : 1, st, 0);
        }
        launch_batch_err<T>(d_films, d_tab + elem_off, elem_n, nfilms, d_partial, d_err, stream);
        LLGMI_HIP(hipMemcpyAsync(h_err, d_err, sizeof(double) * nfilms, hipMemcpyDeviceToHost, stream));
    }
    // `attempts` consecutive attempts of every film with the controller on
    // the device: after each attempt's error norm, kb_control decides and
    // prepares the next attempt and kb_accept copies the accepted states; the
    // film records and the per-attempt outputs come back at the end.
    void launch_run(Run &r, int64_t attempts, double max_err, double headroom, double dt_min, double dt_max) {
        LLGMI_HIP(hipMemcpyAsync(d_films, h_films, sizeof(BFilm<T>) * nfilms, hipMemcpyHostToDevice, stream));
        for (int64_t a = 0; a < attempts; ++a) {
            launch_field(0, 1, 1);
            for (int st = 2; st <= 7; ++st) {
                launch_batch_stage<T>(d_films, d_tab + elem_off, elem_n, st, stream);
                launch_field(st == 7 ? 2 : 1, st, 0);
            }
            launch_batch_err<T>(d_films, d_tab + elem_off, elem_n, nfilms, d_partial, d_err, stream);
            launch_batch_control<T>(d_films, nfilms, d_err, max_err, headroom, dt_min, dt_max, (int)a, r.d_eps,
                                    r.d_acc, r.d_dtu, stream);
            launch_batch_accept<T>(d_films, d_tab + elem_off, elem_n, stream);
        }
        const size_t no = (size_t)(attempts * nfilms);
        LLGMI_HIP(hipMemcpyAsync(h_films, d_films, sizeof(BFilm<T>) * nfilms, hipMemcpyDeviceToHost, stream));
        LLGMI_HIP(hipMemcpyAsync(r.h_eps, r.d_eps, sizeof(double) * no, hipMemcpyDeviceToHost, stream));
        LLGMI_HIP(hipMemcpyAsync(r.h_dtu, r.d_dtu, sizeof(double) * no, hipMemcpyDeviceToHost, stream));
        LLGMI_HIP(hipMemcpyAsync(r.h_acc, r.d_acc, sizeof(int) * no, hipMemcpyDeviceToHost, stream));
    }
};

// the structure a plan depends on: per film its grid, whether the FFT demag
// term is on, and the context's kernel choices
template <typename T>
std::string batch_key(const std::vector<Engine<T> *> &engs, const std::vector<int> &fft_on) {
    std::string k;
    auto put = [&](const void *v, size_t b) { k.append((const char *)v, b); };
    for (size_t f = 0; f < engs.size(); ++f) {
        const Engine<T> &e = *engs[f];
        put(&e.g.nx, sizeof e.g.nx);
        put(&e.g.ny, sizeof e.g.ny);
        put(&fft_on[f], sizeof(int));
        if (fft_on[f]) {
            const int ch[5] = {e.fft.lx, e.fft.ly, e.fft.rq_row, e.fft.rq_col & 7, e.fft.rq_fused > 0 ? e.fft.rq_fused : -e.fft.rq_row_inv};
            put(ch, sizeof ch);
        }
    }
    return k;
}

template <typename T>
std::unique_ptr<BatchPlan<T>> build_batch_plan(const std::vector<Engine<T> *> &engs, const std::vector<int> &fft_on,
                                               int dev) {
    auto P = std::make_unique<BatchPlan<T>>();
    P->dev = dev;
    P->nfilms = (int)engs.size();
    std::vector<BEntry> tab;
    // launch classes of each kind, keyed by (class length, rq, half) / demag
    // flag; a class covers the films of every length whose blocks have the
    // same shape
    std::map<std::tuple<int, int, int, int>, std::vector<BEntry>> cls;
    std::vector<BEntry> elem;
    std::vector<int> e0((size_t)P->nfilms), ne((size_t)P->nfilms);
    for (int f = 0; f < P->nfilms; ++f) {
        const Engine<T> &e = *engs[(size_t)f];
        const int nx = (int)e.g.nx, ny = (int)e.g.ny;
        const int ncell_blocks = (int)((e.n + kBatchCells - 1) / kBatchCells);
        e0[(size_t)f] = (int)elem.size();
        ne[(size_t)f] = ncell_blocks;
        for (int b = 0; b < ncell_blocks; ++b) elem.push_back({f, b});
        if (!fft_on[(size_t)f]) {
            auto &v = cls[{4, 0, 0, 0}];
            for (int b = 0; b < ncell_blocks; ++b) v.push_back({f, b});
            continue;
        }
        const DemagFFT<T> &op = e.fft;
        const bool hx = 2 * op.nx == op.px, hy = 2 * op.ny == op.py;
        {
            auto &v = cls[{0, batch_class_lg<T>(0, op.lx, op.rq_row), op.rq_row, hx}];
            const int nb = batch_row_blocks<T>(0, op.lx, op.rq_row, ny);
            for (int b = 0; b < 3 * nb; ++b) v.push_back({f, b});
        }
        {
            auto &v = cls[{1, batch_class_lg<T>(1, op.ly, op.rq_col & 7), op.rq_col & 7, hy}];
            const int hxs = (int)(op.px / 2), nwork = hxs + (hxs + 1) / 2;
            const int nb = batch_row_blocks<T>(1, op.ly, op.rq_col & 7, nwork);
            for (int w = 0; w < nb; ++w) v.push_back({f, w});
        }
        if (op.rq_fused > 0) {
            auto &v = cls[{3, batch_class_lg<T>(3, op.lx, op.rq_fused), op.rq_fused, hx}];
            const int nb = batch_row_blocks<T>(3, op.lx, op.rq_fused, ny);
            for (int b = 0; b < nb; ++b) v.push_back({f, b});
        } else {
            auto &v = cls[{2, batch_class_lg<T>(2, op.lx, op.rq_row_inv), op.rq_row_inv, hx}];
            const int nb = batch_row_blocks<T>(2, op.lx, op.rq_row_inv, ny);
            for (int b = 0; b < 3 * nb; ++b) v.push_back({f, b});
            auto &w = cls[{5, 0, 0, 0}];   // local field with B_demag, after the inverse passes
            for (int b = 0; b < ncell_blocks; ++b) w.push_back({f, b});
        }
        (void)nx;
    }
    // launch order: all row R2C, all column, all inverse (fused and unfused),
    // then the local fields (the map orders keys by kind); within a class the
    // blocks of the longest transforms go first
    for (auto &kv : cls) {
        const int kind = std::get<0>(kv.first);
        if (kind < 4)
            std::stable_sort(kv.second.begin(), kv.second.end(), [&](const BEntry &a, const BEntry &b) {
                const DemagFFT<T> &fa = engs[(size_t)a.film]->fft, &fb = engs[(size_t)b.film]->fft;
                return kind == 1 ? fa.ly > fb.ly : fa.lx > fb.lx;
            });
    }
    for (auto &kv : cls) {
        const int kind = std::get<0>(kv.first);
        typename BatchPlan<T>::Launch l{};
        l.kind = kind >= 4 ? 4 : kind;
        l.lg = std::get<1>(kv.first);
        l.rq = std::get<2>(kv.first);
        l.half = std::get<3>(kv.first) != 0;
        l.demag = kind == 5 ? 1 : 0;
        l.off = (int)tab.size();
        l.n = (int)kv.second.size();
        tab.insert(tab.end(), kv.second.begin(), kv.second.end());
        P->field.push_back(l);
    }
    P->elem_off = (int)tab.size();
    P->elem_n = (int)elem.size();
    tab.insert(tab.end(), elem.begin(), elem.end());
    for (auto &l : P->field)
        if (l.kind < 4) batch_allow_lds<T>(l.kind, l.lg, l.rq);
    LLGMI_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    LLGMI_HIP(hipMalloc(&P->d_films, sizeof(BFilm<T>) * P->nfilms));
    LLGMI_HIP(hipHostMalloc((void **)&P->h_films, sizeof(BFilm<T>) * P->nfilms, hipHostMallocDefault));
    LLGMI_HIP(hipMalloc(&P->d_tab, sizeof(BEntry) * std::max<size_t>(tab.size(), 1)));
    LLGMI_HIP(hipMemcpy(P->d_tab, tab.data(), sizeof(BEntry) * tab.size(), hipMemcpyHostToDevice));
    LLGMI_HIP(hipMalloc(&P->d_partial, sizeof(double) * std::max(P->elem_n, 1)));
    LLGMI_HIP(hipMalloc(&P->d_err, sizeof(double) * P->nfilms));
    LLGMI_HIP(hipHostMalloc((void **)&P->h_err, sizeof(double) * P->nfilms, hipHostMallocDefault));
    std::memset(P->h_films, 0, sizeof(BFilm<T>) * P->nfilms);
    for (int f = 0; f < P->nfilms; ++f) {
        P->h_films[f].e0 = e0[(size_t)f];
        P->h_films[f].ne = ne[(size_t)f];
    }
    return P;
}

// Process-wide plans, keyed by device, precision and structure (contexts
// that come and go with the same structure reuse a plan: every per-film
// pointer is written into the film records at each attempt).
std::mutex g_batch_mu;
std::map<std::string, std::unique_ptr<BatchPlan<float>>> g_plans_f;
std::map<std::string, std::unique_ptr<BatchPlan<double>>> g_plans_d;
template <typename T>
std::map<std::string, std::unique_ptr<BatchPlan<T>>> &batch_plans();
template <>
std::map<std::string, std::unique_ptr<BatchPlan<float>>> &batch_plans<float>() { return g_plans_f; }
template <>
std::map<std::string, std::unique_ptr<BatchPlan<double>>> &batch_plans<double>() { return g_plans_d; }

bool batch_graphs_enabled() {
    const char *e = std::getenv("LLGMI_BATCH_GRAPH");
    return !(e && e[0] == '0');
}

// The records of every film for an attempt (after rk45_prepare).
template <typename T>
void fill_film_records(BatchPlan<T> &P, const std::vector<Engine<T> *> &engs, const std::vector<T *> &ms,
                       const llgmi_params *params, const llgmi_material *mats, int mask,
                       const std::vector<int> &fft_on, const std::vector<int> &fsal) {
    const int nf = (int)engs.size();
    for (int f = 0; f < nf; ++f) {
        Engine<T> &e = *engs[(size_t)f];
        BFilm<T> &r = P.h_films[f];
        const int e0 = r.e0, ne = r.ne;
        r = BFilm<T>{};
        r.e0 = e0;
        r.ne = ne;
        r.nx = (int)e.g.nx;
        r.ny = (int)e.g.ny;
        r.fft = fft_on[(size_t)f];
        if (r.fft) {
            r.spec = e.fft.spec;
            r.tw_x = e.fft.tw_x;
            r.tw_y = e.fft.tw_y;
            r.kern = e.fft.kern;
            r.px = (int)e.fft.px;
            r.py = (int)e.fft.py;
        }
        r.ms = T(mats[f].ms);
        r.skip1 = fsal[(size_t)f];
        r.dt = params[f].dt;
        r.la = e.local_args(params[f], mats[f], mask, true, false);
        r.m = ms[(size_t)f];
        r.in[0] = ms[(size_t)f];
        r.in[1] = e.d_stage;
        r.in[2] = e.d_m7;
        r.stage = e.d_stage;
        r.m7 = e.d_m7;
        r.b = nullptr;   // the stages keep only the torque
        r.bd = e.d_bd;
        for (int q = 0; q < 8; ++q) r.k[q] = e.d_k[q];
    }
}

// Longest device-controlled run (attempts per call of llgmi_dev_rk45_run*).
constexpr int64_t kRunAttemptsMax = 256;

// One batched attempt of every film (the films' engines on one device, all
// with their operators built and tuned).  eps / accepted / dt_used per film.
template <typename T>
void batch_attempt(const std::vector<Engine<T> *> &engs, const std::vector<T *> &ms, llgmi_params *params,
                   const llgmi_material *mats, int mask, double max_err, double headroom, double dt_min,
                   double dt_max, double *eps, int *accepted, double *dt_used, int dev) {
    const int nf = (int)engs.size();
    std::vector<int> fft_on((size_t)nf);
    for (int f = 0; f < nf; ++f)
        fft_on[(size_t)f] = mask == LLGMI_MASK_FULL && mats[f].demag && resolved_demag_method(mats[f]) == LLGMI_DEMAG_FFT;
    // every film's earlier work on its own stream is done before the batch stream touches its buffers
    for (auto *e : engs) LLGMI_HIP(hipStreamSynchronize(e->stream));
    std::lock_guard<std::mutex> lock(g_batch_mu);
    std::string key = batch_key<T>(engs, fft_on);
    key.append((const char *)&dev, sizeof dev);
    auto &plans = batch_plans<T>();
    auto it = plans.find(key);
    if (it == plans.end()) {
        if (plans.size() >= 8) plans.clear();
        it = plans.emplace(key, build_batch_plan<T>(engs, fft_on, dev)).first;
    }
    BatchPlan<T> &P = *it->second;
    std::vector<int> fsal((size_t)nf);
    for (int f = 0; f < nf; ++f) fsal[(size_t)f] = engs[(size_t)f]->rk45_prepare(params[f], mats[f], mask) ? 1 : 0;
    fill_film_records<T>(P, engs, ms, params, mats, mask, fft_on, fsal);
    try {
        if (batch_graphs_enabled()) {
            if (!P.exec) {
                hipGraph_t graph = nullptr;
                LLGMI_HIP(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
                try {
                    P.launch_attempt();
                } catch (...) {
                    (void)hipStreamEndCapture(P.stream, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                LLGMI_HIP(hipStreamEndCapture(P.stream, &graph));
                const hipError_t ge = hipGraphInstantiate(&P.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                LLGMI_HIP(ge);
            }
            LLGMI_HIP(hipGraphLaunch(P.exec, P.stream));
        } else {
            P.launch_attempt();
        }
        LLGMI_HIP(hipStreamSynchronize(P.stream));
    } catch (...) {
        for (auto *e : engs) e->issued_mask = -1;   // no film takes this attempt
        throw;
    }
    bool any = false;
    for (int f = 0; f < nf; ++f) {
        engs[(size_t)f]->rk45_decide(ms[(size_t)f], params[f], mask, max_err, headroom, dt_min, dt_max,
                                     P.h_err[f], false, eps[f], accepted[f], dt_used[f]);
        P.h_films[f].accept = accepted[f];
        any = any || accepted[f];
    }
    if (any) {
        LLGMI_HIP(hipMemcpyAsync(P.d_films, P.h_films, sizeof(BFilm<T>) * nf, hipMemcpyHostToDevice, P.stream));
        launch_batch_accept<T>(P.d_films, P.d_tab + P.elem_off, P.elem_n, P.stream);
        LLGMI_HIP(hipStreamSynchronize(P.stream));
    }
}

// `attempts` consecutive adaptive attempts of every film with the controller
// on the device (batched kernels, one graph replay, one host synchronisation
// for all).  Outputs are attempt-major: [a * films + f].  The films' engines
// take the state the last attempt leaves (step size, FSAL, k1/k7).
template <typename T>
void batch_run(const std::vector<Engine<T> *> &engs, const std::vector<T *> &ms, llgmi_params *params,
               const llgmi_material *mats, int mask, double max_err, double headroom, double dt_min, double dt_max,
               int64_t attempts, double *eps, int *accepted, double *dt_used, int dev) {
    const int nf = (int)engs.size();
    std::vector<int> fft_on((size_t)nf);
    for (int f = 0; f < nf; ++f)
        fft_on[(size_t)f] = mask == LLGMI_MASK_FULL && mats[f].demag && resolved_demag_method(mats[f]) == LLGMI_DEMAG_FFT;
    for (auto *e : engs) LLGMI_HIP(hipStreamSynchronize(e->stream));
    std::lock_guard<std::mutex> lock(g_batch_mu);
    std::string key = batch_key<T>(engs, fft_on);
    key.append((const char *)&dev, sizeof dev);
    auto &plans = batch_plans<T>();
    auto it = plans.find(key);
    if (it == plans.end()) {
        if (plans.size() >= 8) plans.clear();
        it = plans.emplace(key, build_batch_plan<T>(engs, fft_on, dev)).first;
    }
    BatchPlan<T> &P = *it->second;
    std::vector<int> fsal((size_t)nf);
    for (int f = 0; f < nf; ++f) fsal[(size_t)f] = engs[(size_t)f]->rk45_prepare(params[f], mats[f], mask) ? 1 : 0;
    fill_film_records<T>(P, engs, ms, params, mats, mask, fft_on, fsal);
    const auto rkey = std::make_tuple(attempts, max_err, headroom, dt_min, dt_max);
    try {
        auto rit = P.runs.find(rkey);
        if (rit == P.runs.end()) {
            if (P.runs.size() >= 4) {
                for (auto &kv : P.runs) {
                    typename BatchPlan<T>::Run &r = kv.second;
                    if (r.exec) (void)hipGraphExecDestroy(r.exec);
                    void *dv[] = {r.d_eps, r.d_dtu, r.d_acc};
                    for (void *p : dv)
                        if (p) (void)hipFree(p);
                    void *hv[] = {r.h_eps, r.h_dtu, r.h_acc};
                    for (void *p : hv)
                        if (p) (void)hipHostFree(p);
                }
                P.runs.clear();
            }
            rit = P.runs.emplace(rkey, typename BatchPlan<T>::Run{}).first;
            typename BatchPlan<T>::Run &r = rit->second;
            const size_t no = (size_t)(attempts * nf);
            LLGMI_HIP(hipMalloc(&r.d_eps, sizeof(double) * no));
            LLGMI_HIP(hipMalloc(&r.d_dtu, sizeof(double) * no));
            LLGMI_HIP(hipMalloc(&r.d_acc, sizeof(int) * no));
            LLGMI_HIP(hipHostMalloc((void **)&r.h_eps, sizeof(double) * no, hipHostMallocDefault));
            LLGMI_HIP(hipHostMalloc((void **)&r.h_dtu, sizeof(double) * no, hipHostMallocDefault));
            LLGMI_HIP(hipHostMalloc((void **)&r.h_acc, sizeof(int) * no, hipHostMallocDefault));
        }
        typename BatchPlan<T>::Run &r = rit->second;
        if (batch_graphs_enabled()) {
            if (!r.exec) {
                hipGraph_t graph = nullptr;
                LLGMI_HIP(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
                try {
                    P.launch_run(r, attempts, max_err, headroom, dt_min, dt_max);
                } catch (...) {
                    (void)hipStreamEndCapture(P.stream, &graph);
                    if (graph) (void)hipGraphDestroy(graph);
                    throw;
                }
                LLGMI_HIP(hipStreamEndCapture(P.stream, &graph));
                const hipError_t ge = hipGraphInstantiate(&r.exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                LLGMI_HIP(ge);
            }
            LLGMI_HIP(hipGraphLaunch(r.exec, P.stream));
        } else {
            P.launch_run(r, attempts, max_err, headroom, dt_min, dt_max);
        }
        LLGMI_HIP(hipStreamSynchronize(P.stream));
        const size_t no = (size_t)(attempts * nf);
        std::memcpy(eps, r.h_eps, sizeof(double) * no);
        std::memcpy(dt_used, r.h_dtu, sizeof(double) * no);
        std::memcpy(accepted, r.h_acc, sizeof(int) * no);
    } catch (...) {
        // a failed run leaves the films' states undefined: no FSAL reuse
        for (auto *e : engs) {
            e->issued_mask = -1;
            e->fsal_valid = false;
            e->last_mask = -1;
        }
        throw;
    }
    for (int f = 0; f < nf; ++f) {
        Engine<T> &e = *engs[(size_t)f];
        const BFilm<T> &r = P.h_films[f];
        const bool acc = r.accept != 0;
        // The device swaps k1 and k7 right after an accepted attempt, the host
        // path at the start of the next attempt (rk45_prepare): after an
        // accepted last attempt the engine keeps the pointers from before
        // that swap, and its next attempt swaps them itself.
        T *k1 = acc ? r.k[7] : r.k[1];
        if (k1 != e.d_k[1]) {
            std::swap(e.d_k[1], e.d_k[7]);
            e.k_swapped = !e.k_swapped;
        }
        e.fsal_valid = acc;
        e.last_mask = acc ? mask : -1;
        if (acc)
            for (int q = 0; q < 3; ++q) e.last_bext[q] = params[f].b_ext[q];
        e.issued_mask = -1;
        params[f].dt = r.dt;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *llgmi_last_error(void) { return g_last_error.c_str(); }
int llgmi_version(void) { return 1; }

int llgmi_device_count(int *count) {
    if (!count) return fail(LLGMI_ERR_INVALID, "count is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        c = 0;
    }
    *count = c;
    return LLGMI_OK;
}

int llgmi_ctx_create(const llgmi_grid *grid, int precision, int device, llgmi_ctx **out) {
    if (!out) return fail(LLGMI_ERR_INVALID, "out is NULL");
    *out = nullptr;
    std::string e = check_grid(grid);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    if (precision != LLGMI_F32 && precision != LLGMI_F64) return fail(LLGMI_ERR_INVALID, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(LLGMI_ERR_HIP, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(LLGMI_ERR_INVALID, "device index out of range");
    std::unique_ptr<llgmi_ctx> c(new (std::nothrow) llgmi_ctx());
    if (!c) return fail(LLGMI_ERR_ALLOC, "host allocation failed");
    c->grid = *grid;
    c->precision = precision;
    c->device = device;
    try {
        LLGMI_HIP(hipSetDevice(device));
        ensure_scratch_address(device);
        LLGMI_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        if (precision == LLGMI_F32) c->ef.reset(new Engine<float>(*grid, c->stream, &c->prof));
        else c->ed.reset(new Engine<double>(*grid, c->stream, &c->prof));
    } catch (const Error &ex) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        return fail(ex.code, ex.what());
    }
    *out = c.release();
    return LLGMI_OK;
}

int llgmi_ctx_destroy(llgmi_ctx *ctx) {
    if (!ctx) return LLGMI_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->ef.reset();
    ctx->ed.reset();
    ctx->prof.pending.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LLGMI_OK;
}

int llgmi_ctx_stream(llgmi_ctx *ctx, void **s) {
    if (!ctx || !s) return fail(LLGMI_ERR_INVALID, "NULL argument");
    *s = (void *)ctx->stream;
    return LLGMI_OK;
}

int llgmi_ctx_synchronize(llgmi_ctx *ctx) {
    return guarded(ctx, [&] { LLGMI_HIP(hipStreamSynchronize(ctx->stream)); });
}

int llgmi_build_h_eff_masked_geom(llgmi_ctx *ctx, const double *m, double *b_eff,
                                  const llgmi_params *params, const llgmi_material *mat,
                                  int field_mask, const uint8_t *geom_mask) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m || !b_eff || !params) return fail(LLGMI_ERR_INVALID, "NULL field or params");
    std::string e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) { eng.host_build(m, b_eff, *params, *mat, field_mask, geom_mask); });
    });
}

int llgmi_build_h_eff_masked(llgmi_ctx *ctx, const double *m, double *b_eff,
                             const llgmi_params *params, const llgmi_material *mat,
                             int field_mask) {
    return llgmi_build_h_eff_masked_geom(ctx, m, b_eff, params, mat, field_mask, nullptr);
}

int llgmi_build_h_eff(llgmi_ctx *ctx, const double *m, double *b_eff, const llgmi_params *params,
                      const llgmi_material *mat) {
    return llgmi_build_h_eff_masked_geom(ctx, m, b_eff, params, mat, LLGMI_MASK_FULL, nullptr);
}

int llgmi_add_demag_field_pbc(llgmi_ctx *ctx, const double *m, double *b_eff,
                              const llgmi_material *mat, uint64_t pbc_x, uint64_t pbc_y) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m || !b_eff) return fail(LLGMI_ERR_INVALID, "NULL field");
    std::string e = check_mat(mat);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) { eng.host_add_demag(m, b_eff, *mat, pbc_x, pbc_y); });
    });
}

int llgmi_compute_demag_field_pbc(llgmi_ctx *ctx, const double *m, double *out,
                                  const llgmi_material *mat, uint64_t pbc_x, uint64_t pbc_y) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!out) return fail(LLGMI_ERR_INVALID, "NULL field");
    std::memset(out, 0, sizeof(double) * 3 * (size_t)(ctx->grid.nx * ctx->grid.ny));
    return llgmi_add_demag_field_pbc(ctx, m, out, mat, pbc_x, pbc_y);
}

int llgmi_add_demag_field(llgmi_ctx *ctx, const double *m, double *b_eff, const llgmi_material *mat) {
    return llgmi_add_demag_field_pbc(ctx, m, b_eff, mat, 0, 0);
}

int llgmi_rk45_step(llgmi_ctx *ctx, double *m, llgmi_params *params, const llgmi_material *mat,
                    int field_mask, double max_err, double headroom, double dt_min, double dt_max,
                    double *eps, int *accepted, double *dt_used) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m || !params || !eps || !accepted || !dt_used) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    if (dt_min > dt_max) return fail(LLGMI_ERR_INVALID, "dt_min > dt_max");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            eng.ensure_io();
            eng.upload(m, eng.d_m);
            eng.rk45(eng.d_m, *params, *mat, field_mask, max_err, headroom, dt_min, dt_max, *eps,
                     *accepted, *dt_used);
            if (*accepted) eng.download(eng.d_m, m);
        });
    });
}

int llgmi_rk45_step_adaptive(llgmi_ctx *ctx, double *m, llgmi_params *params,
                             const llgmi_material *mat, double max_err, double headroom,
                             double dt_min, double dt_max, double *eps, int *accepted,
                             double *dt_used) {
    return llgmi_rk45_step(ctx, m, params, mat, LLGMI_MASK_FULL, max_err, headroom, dt_min, dt_max,
                           eps, accepted, dt_used);
}

int llgmi_rk45_reset(llgmi_ctx *ctx) {
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            eng.fsal_valid = false;
            eng.last_mask = -1;
        });
    });
}

int llgmi_dev_build_h_eff(llgmi_ctx *ctx, const void *m_dev, void *b_dev, const llgmi_params *params,
                          const llgmi_material *mat, int field_mask, const uint8_t *geom_mask_dev) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!params) return fail(LLGMI_ERR_INVALID, "params is NULL");
    std::string e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (e.empty()) e = check_dev(b_dev, 3 * n * elem_size(ctx), "b");
    if (e.empty() && geom_mask_dev) e = check_dev(geom_mask_dev, n, "geom_mask");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.field((const T *)m_dev, (T *)b_dev, *params, *mat, field_mask, geom_mask_dev, nullptr);
        });
    });
}

int llgmi_dev_build_h_eff_noalloc(llgmi_ctx *ctx, const void *m_dev, void *b_dev, const llgmi_params *params,
                                  const llgmi_material *mat, int field_mask, const uint8_t *geom_mask_dev) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!params) return fail(LLGMI_ERR_INVALID, "params is NULL");
    std::string e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (e.empty()) e = check_dev(b_dev, 3 * n * elem_size(ctx), "b");
    if (e.empty() && geom_mask_dev) e = check_dev(geom_mask_dev, n, "geom_mask");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    void *scr = scratch_address(ctx->device);
    if (!scr) return fail(LLGMI_ERR_INVALID, "no scratch address for the context's device");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.field_noalloc((const T *)m_dev, (T *)b_dev, *params, *mat, field_mask, geom_mask_dev,
                              (unsigned char *)scr, kNoAllocScratchBytes);
        });
    });
}

int llgmi_dev_add_demag_field_pbc(llgmi_ctx *ctx, const void *m_dev, void *b_dev,
                                  const llgmi_material *mat, uint64_t pbc_x, uint64_t pbc_y) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    std::string e = check_mat(mat);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (e.empty()) e = check_dev(b_dev, 3 * n * elem_size(ctx), "b");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    if (!mat->demag) return LLGMI_OK;
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.demag_into_bd((const T *)m_dev, *mat, resolved_demag_method(*mat), pbc_x, pbc_y);
            launch_soa_accumulate<T>(eng.d_bd, (T *)b_dev, 3 * eng.n, eng.stream);
        });
    });
}

int llgmi_dev_rk45_step(llgmi_ctx *ctx, void *m_dev, llgmi_params *params, const llgmi_material *mat,
                        int field_mask, double max_err, double headroom, double dt_min,
                        double dt_max, double *eps, int *accepted, double *dt_used) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!params || !eps || !accepted || !dt_used) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    if (dt_min > dt_max) return fail(LLGMI_ERR_INVALID, "dt_min > dt_max");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.rk45((T *)m_dev, *params, *mat, field_mask, max_err, headroom, dt_min, dt_max, *eps,
                     *accepted, *dt_used);
        });
    });
}

int llgmi_dev_rk45_step_batch(llgmi_ctx *const *ctxs, void *const *m_devs, llgmi_params *params,
                              const llgmi_material *mats, int64_t count, int field_mask, double max_err,
                              double headroom, double dt_min, double dt_max, double *eps, int *accepted,
                              double *dt_used) {
    if (count < 0) return fail(LLGMI_ERR_INVALID, "count < 0");
    if (count == 0) return LLGMI_OK;
    if (!ctxs || !m_devs || !params || !mats || !eps || !accepted || !dt_used)
        return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mask(field_mask);
    if (e.empty() && dt_min > dt_max) e = "dt_min > dt_max";
    std::set<const llgmi_ctx *> seen;
    for (int64_t i = 0; i < count && e.empty(); ++i) {
        const llgmi_ctx *c = ctxs[i];
        if (!c) { e = "context is NULL"; break; }
        if (!seen.insert(c).second) { e = "a context appears twice in one batch"; break; }
        e = check_mat(&mats[i]);
        if (e.empty()) {
            (void)hipSetDevice(c->device);
            e = check_dev(m_devs[i], 3 * (size_t)(c->grid.nx * c->grid.ny) * elem_size(c), "m");
        }
        if (!e.empty()) e = "film " + std::to_string(i) + ": " + e;
    }
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    // Batched kernels once every film's operator is built and tuned (the
    // first call for a set of films goes through the per-film path, which
    // does that), for films of one precision on one device, the FFT or no
    // demag term, profiling off; LLGMI_BATCH_KERNELS=0 keeps the per-film path.
    {
        const char *bk = std::getenv("LLGMI_BATCH_KERNELS");
        bool ok = count >= 2 && !(bk && bk[0] == '0');
        for (int64_t i = 0; i < count && ok; ++i) {
            const llgmi_ctx *c = ctxs[i];
            ok = c->precision == ctxs[0]->precision && c->device == ctxs[0]->device && c->prof.mask == 0;
            if (!ok) break;
            const bool demag = field_mask == LLGMI_MASK_FULL && mats[i].demag;
            if (demag && resolved_demag_method(mats[i]) != LLGMI_DEMAG_FFT) { ok = false; break; }
            auto ready = [&](auto &eng) {
                if (!demag) return true;
                // split column passes (FP64 columns of 2^13) and 3 * 2^k lengths are not
                // in the batched kernels
                return eng.fft.built && eng.fft.rq_fused != 0 && eng.fft.colfull == nullptr && !eng.fft.has_f3();
            };
            ok = c->precision == LLGMI_F32 ? ready(*c->ef) : ready(*c->ed);
        }
        if (ok) {
            llgmi_ctx *c0 = ctxs[0];
            return guarded(c0, [&] {
                if (c0->precision == LLGMI_F32) {
                    std::vector<Engine<float> *> engs;
                    std::vector<float *> ms;
                    for (int64_t i = 0; i < count; ++i) {
                        engs.push_back(ctxs[i]->ef.get());
                        ms.push_back((float *)m_devs[i]);
                    }
                    batch_attempt<float>(engs, ms, params, mats, field_mask, max_err, headroom, dt_min, dt_max, eps,
                                         accepted, dt_used, c0->device);
                } else {
                    std::vector<Engine<double> *> engs;
                    std::vector<double *> ms;
                    for (int64_t i = 0; i < count; ++i) {
                        engs.push_back(ctxs[i]->ed.get());
                        ms.push_back((double *)m_devs[i]);
                    }
                    batch_attempt<double>(engs, ms, params, mats, field_mask, max_err, headroom, dt_min, dt_max,
                                          eps, accepted, dt_used, c0->device);
                }
            });
        }
    }
    // every attempt is launched before any is waited for
    int64_t issued = 0;
    int rc = LLGMI_OK;
    for (; issued < count; ++issued) {
        llgmi_ctx *c = ctxs[issued];
        rc = guarded(c, [&] {
            with_engine(c, [&](auto &eng) {
                using T = std::remove_pointer_t<decltype(eng.d_m)>;
                eng.rk45_issue((T *)m_devs[issued], params[issued], mats[issued], field_mask);
            });
        });
        if (rc != LLGMI_OK) break;
    }
    std::string first_err = rc != LLGMI_OK ? llgmi_last_error() : "";
    for (int64_t i = 0; i < issued; ++i) {
        llgmi_ctx *c = ctxs[i];
        const bool finish = rc == LLGMI_OK;
        const int r = guarded(c, [&] {
            with_engine(c, [&](auto &eng) {
                using T = std::remove_pointer_t<decltype(eng.d_m)>;
                LLGMI_HIP(hipStreamSynchronize(eng.stream));
                if (finish)
                    eng.rk45_finish((T *)m_devs[i], params[i], field_mask, max_err, headroom, dt_min, dt_max,
                                    eps[i], accepted[i], dt_used[i]);
                else
                    eng.issued_mask = -1;   // the batch failed: no film takes this attempt
            });
        });
        if (r != LLGMI_OK && rc == LLGMI_OK) {
            rc = r;
            first_err = llgmi_last_error();
        }
    }
    if (rc != LLGMI_OK) return fail(rc, first_err);
    return LLGMI_OK;
}

int llgmi_dev_rk45_run_batch(llgmi_ctx *const *ctxs, void *const *m_devs, llgmi_params *params,
                             const llgmi_material *mats, int64_t count, int64_t attempts, int field_mask,
                             double max_err, double headroom, double dt_min, double dt_max, double *eps,
                             int *accepted, double *dt_used) {
    if (count < 0) return fail(LLGMI_ERR_INVALID, "count < 0");
    if (attempts < 0 || attempts > kRunAttemptsMax)
        return fail(LLGMI_ERR_INVALID, "attempts must be in [0, " + std::to_string(kRunAttemptsMax) + "]");
    if (count == 0 || attempts == 0) return LLGMI_OK;
    if (!ctxs || !m_devs || !params || !mats || !eps || !accepted || !dt_used)
        return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mask(field_mask);
    if (e.empty() && dt_min > dt_max) e = "dt_min > dt_max";
    std::set<const llgmi_ctx *> seen;
    for (int64_t i = 0; i < count && e.empty(); ++i) {
        const llgmi_ctx *c = ctxs[i];
        if (!c) { e = "context is NULL"; break; }
        if (!seen.insert(c).second) { e = "a context appears twice in one batch"; break; }
        e = check_mat(&mats[i]);
        if (e.empty()) {
            (void)hipSetDevice(c->device);
            e = check_dev(m_devs[i], 3 * (size_t)(c->grid.nx * c->grid.ny) * elem_size(c), "m");
        }
        if (!e.empty()) e = "film " + std::to_string(i) + ": " + e;
    }
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    for (int64_t i = 0; i < count; ++i) {
        const llgmi_ctx *c = ctxs[i];
        std::string why;
        if (c->precision != ctxs[0]->precision || c->device != ctxs[0]->device)
            why = "films of one precision on one device";
        else if (c->prof.mask != 0)
            why = "kernel profiling off";
        else if (field_mask == LLGMI_MASK_FULL && mats[i].demag && resolved_demag_method(mats[i]) != LLGMI_DEMAG_FFT)
            why = "the FFT demag (or none)";
        if (!why.empty())
            return fail(LLGMI_ERR_UNSUPPORTED, "film " + std::to_string(i) + ": device-controlled runs need " + why);
    }
    if (count == 1)
        return llgmi_dev_rk45_run(ctxs[0], m_devs[0], params, mats, attempts, field_mask, max_err, headroom, dt_min,
                                  dt_max, eps, accepted, dt_used);
    llgmi_ctx *c0 = ctxs[0];
    return guarded(c0, [&] {
        auto run = [&](auto tag, auto getter) {
            using T = decltype(tag);
            std::vector<Engine<T> *> engs;
            std::vector<T *> ms;
            for (int64_t i = 0; i < count; ++i) {
                Engine<T> &eng = *getter(ctxs[i]);
                const bool demag = field_mask == LLGMI_MASK_FULL && mats[i].demag;
                if (demag && !(eng.fft.built && eng.fft.rq_fused != 0)) {
                    // builds and tunes the film's operator (the batched kernels use its choices)
                    eng.ensure_rk();
                    eng.field((const T *)m_devs[i], eng.d_b, params[i], mats[i], field_mask, nullptr, nullptr);
                    LLGMI_HIP(hipStreamSynchronize(eng.stream));
                }
                if (demag && (eng.fft.colfull != nullptr || eng.fft.has_f3()))
                    throw Error(LLGMI_ERR_UNSUPPORTED,
                                "film " + std::to_string(i) +
                                    ": split column passes (FP64, ny > 2048) and padded lengths 3 * 2^k are not "
                                    "batched; step such films with llgmi_dev_rk45_run one at a time");
                engs.push_back(&eng);
                ms.push_back((T *)m_devs[i]);
            }
            batch_run<T>(engs, ms, params, mats, field_mask, max_err, headroom, dt_min, dt_max, attempts, eps,
                         accepted, dt_used, c0->device);
        };
        if (c0->precision == LLGMI_F32)
            run(float{}, [](llgmi_ctx *c) { return c->ef.get(); });
        else
            run(double{}, [](llgmi_ctx *c) { return c->ed.get(); });
    });
}

int llgmi_dev_rk45_run(llgmi_ctx *ctx, void *m_dev, llgmi_params *params, const llgmi_material *mat,
                       int64_t attempts, int field_mask, double max_err, double headroom, double dt_min,
                       double dt_max, double *eps, int *accepted, double *dt_used) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (attempts < 0 || attempts > kRunAttemptsMax)
        return fail(LLGMI_ERR_INVALID, "attempts must be in [0, " + std::to_string(kRunAttemptsMax) + "]");
    if (attempts == 0) return LLGMI_OK;
    if (!params || !mat || !eps || !accepted || !dt_used) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mask(field_mask);
    if (e.empty()) e = check_mat(mat);
    if (e.empty() && dt_min > dt_max) e = "dt_min > dt_max";
    if (e.empty()) e = check_dev(m_dev, 3 * (size_t)(ctx->grid.nx * ctx->grid.ny) * elem_size(ctx), "m");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.rk45_run((T *)m_dev, *params, *mat, field_mask, max_err, headroom, dt_min, dt_max, attempts, eps,
                         accepted, dt_used);
        });
    });
}

int llgmi_dev_mask_cells(llgmi_ctx *ctx, void *m_dev, const uint8_t *keep_dev) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    std::string e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (e.empty()) e = check_dev(keep_dev, n, "keep");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            launch_mask_cells<T>((T *)m_dev, keep_dev, eng.n, eng.stream);
            LLGMI_HIP(hipStreamSynchronize(eng.stream));
        });
    });
}

int llgmi_dev_upload(llgmi_ctx *ctx, const double *host_aos, void *dev_soa) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!host_aos) return fail(LLGMI_ERR_INVALID, "host buffer is NULL");
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    std::string e = check_dev(dev_soa, 3 * n * elem_size(ctx), "dev_soa");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.ensure_io();
            eng.upload(host_aos, (T *)dev_soa);
            LLGMI_HIP(hipStreamSynchronize(eng.stream));
        });
    });
}

int llgmi_dev_download(llgmi_ctx *ctx, const void *dev_soa, double *host_aos) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!host_aos) return fail(LLGMI_ERR_INVALID, "host buffer is NULL");
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    std::string e = check_dev(dev_soa, 3 * n * elem_size(ctx), "dev_soa");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.ensure_io();
            eng.download((const T *)dev_soa, host_aos);
        });
    });
}

int llgmi_mg_config_default(llgmi_mg_config *cfg) {
    if (!cfg) return fail(LLGMI_ERR_INVALID, "cfg is NULL");
    mg_config_default(*cfg);
    return LLGMI_OK;
}

int llgmi_mg_config_from_env(llgmi_mg_config *cfg) {
    if (!cfg) return fail(LLGMI_ERR_INVALID, "cfg is NULL");
    mg_config_from_env(*cfg);
    return LLGMI_OK;
}

int llgmi_ctx_set_mg_config(llgmi_ctx *ctx, const llgmi_mg_config *cfg) {
    return guarded(ctx, [&] {
        if (cfg) mg_validate(*cfg);
        with_engine(ctx, [&](auto &eng) {
            eng.has_mg_cfg = cfg != nullptr;
            if (cfg) eng.mg_cfg = *cfg;
        });
    });
}

int llgmi_add_demag_field_poisson_mg(llgmi_ctx *ctx, const double *m, double *b_eff,
                                     const llgmi_material *mat) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m || !b_eff) return fail(LLGMI_ERR_INVALID, "NULL field");
    std::string e = check_mat(mat);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) { eng.host_add_demag(m, b_eff, *mat, 0, 0, LLGMI_DEMAG_MG); });
    });
}

int llgmi_compute_demag_field_poisson_mg(llgmi_ctx *ctx, const double *m, double *out,
                                         const llgmi_material *mat) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!out) return fail(LLGMI_ERR_INVALID, "NULL field");
    std::memset(out, 0, sizeof(double) * 3 * (size_t)(ctx->grid.nx * ctx->grid.ny));
    return llgmi_add_demag_field_poisson_mg(ctx, m, out, mat);
}

int llgmi_mg_hybrid_config_default(llgmi_mg_hybrid_config *cfg) {
    if (!cfg) return fail(LLGMI_ERR_INVALID, "config is NULL");
    mg_hybrid_config_default(*cfg);
    return LLGMI_OK;
}

int llgmi_mg_hybrid_config_from_env(llgmi_mg_hybrid_config *cfg) {
    if (!cfg) return fail(LLGMI_ERR_INVALID, "config is NULL");
    mg_hybrid_config_from_env(*cfg);
    return LLGMI_OK;
}

int llgmi_ctx_set_mg_hybrid_config(llgmi_ctx *ctx, const llgmi_mg_hybrid_config *cfg) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (cfg && (cfg->radius_xy < 0 || cfg->delta_v_cycles < 1 || !(cfg->sigma_cells >= 0.0) ||
                cfg->sigma_cells > 32.0 || cfg->bmargin < 0 || cfg->radius_xy > 4096))
        return fail(LLGMI_ERR_INVALID, "hybrid config out of range (radius 0..4096, delta_v_cycles >= 1, "
                                       "sigma 0..32, bmargin >= 0)");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            eng.has_hyb_cfg = cfg != nullptr;
            if (cfg) eng.hyb_cfg = *cfg;
        });
    });
}

int llgmi_mg_last_stats(llgmi_ctx *ctx, llgmi_mg_stats *out) {
    if (!out) return fail(LLGMI_ERR_INVALID, "out is NULL");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            if (!eng.mg) throw Error(LLGMI_ERR_INVALID, "no multigrid solve has run on this context");
            *out = eng.mg->stats();
        });
    });
}

int llgmi_mg_box(llgmi_ctx *ctx, int64_t dims[6]) {
    if (!dims) return fail(LLGMI_ERR_INVALID, "dims is NULL");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            if (!eng.mg) throw Error(LLGMI_ERR_INVALID, "no multigrid solve has run on this context");
            eng.mg->box(dims);
        });
    });
}

int llgmi_mg_phi_read(llgmi_ctx *ctx, double *phi_dev) {
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            if (!eng.mg) throw Error(LLGMI_ERR_INVALID, "no multigrid solve has run on this context");
            int64_t d[6];
            eng.mg->box(d);
            const std::string e = check_dev(phi_dev, sizeof(double) * (size_t)(d[0] * d[1] * d[2]), "phi");
            if (!e.empty()) throw Error(LLGMI_ERR_INVALID, e);
            eng.mg->phi_copy(phi_dev, false);
        });
    });
}

int llgmi_mg_phi_write(llgmi_ctx *ctx, const double *phi_dev) {
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            if (!eng.mg) throw Error(LLGMI_ERR_INVALID, "no multigrid solve has run on this context");
            int64_t d[6];
            eng.mg->box(d);
            const std::string e = check_dev(phi_dev, sizeof(double) * (size_t)(d[0] * d[1] * d[2]), "phi");
            if (!e.empty()) throw Error(LLGMI_ERR_INVALID, e);
            eng.mg->phi_copy(const_cast<double *>(phi_dev), true);
        });
    });
}

// ---- relaxation ----------------------------------------------------------------
namespace {

void relax_settings_default(llgmi_relax_settings &s) {
    std::memset(&s, 0, sizeof s);
    s.max_err = 1e-5;
    s.headroom = 0.8;
    s.dt_min = 1e-18;
    s.dt_max = 1e-11;
    s.phase1_enabled = 1;
    s.phase2_enabled = 1;
    s.energy_stride = 3;
    s.rel_energy_tol = 1e-12;
    s.torque_metric = LLGMI_TORQUE_MAX;
    s.has_torque_threshold = 1;
    s.torque_threshold = 1e-4;
    s.torque_check_stride = 1;
    s.torque_plateau_rel = 1e-3;
    s.torque_plateau_abs = 0.0;
    s.torque_plateau_checks = 0;
    s.torque_plateau_min_checks = 5;
    s.tighten_factor = 0.70710678118654752440;   // FRAC_1_SQRT_2
    s.tighten_floor = 1e-9;
    s.max_accepted_steps = 2000000;
}

std::string check_relax_settings(const llgmi_relax_settings *s) {
    if (!s) return "settings is NULL";
    if (s->torque_metric < 0 || s->torque_metric > 3) return "unknown torque_metric";
    if (s->dt_min > s->dt_max) return "dt_min > dt_max";
    if (s->max_accepted_steps < 0) return "max_accepted_steps must be >= 0";
    if (!(s->tighten_factor > 0.0 && s->tighten_factor < 1.0) && s->tighten_floor > 0.0 &&
        s->max_err > s->tighten_floor)
        return "tighten_factor must lie in (0, 1) for max_err to reach tighten_floor";
    if (!(s->tighten_floor > 0.0) && s->phase2_enabled)
        return "tighten_floor must be positive";
    return "";
}

std::string check_step_args(const llgmi_params *params, const double *eps, const int *accepted,
                            const double *dt_used, double dt_min, double dt_max) {
    if (!params || !eps || !accepted || !dt_used) return "NULL argument";
    if (dt_min > dt_max) return "dt_min > dt_max";
    return "";
}

}  // namespace

int llgmi_relax_settings_default(llgmi_relax_settings *s) {
    if (!s) return fail(LLGMI_ERR_INVALID, "settings is NULL");
    relax_settings_default(*s);
    return LLGMI_OK;
}

int llgmi_relax_settings_mumax_like(llgmi_relax_settings *s) {
    if (!s) return fail(LLGMI_ERR_INVALID, "settings is NULL");
    relax_settings_default(*s);
    s->phase1_enabled = 1;
    s->phase2_enabled = 1;
    s->energy_stride = 3;
    s->rel_energy_tol = 0.0;
    s->torque_metric = LLGMI_TORQUE_MEANSQ;
    s->has_torque_threshold = 0;
    s->torque_threshold = 0.0;
    s->torque_plateau_checks = 1;
    s->torque_plateau_min_checks = 1;
    s->torque_plateau_rel = 0.0;
    s->torque_plateau_abs = 0.0;
    s->tighten_factor = 0.70710678118654752440;
    s->tighten_floor = 1e-9;
    return LLGMI_OK;
}

int llgmi_rk23_relax_step(llgmi_ctx *ctx, double *m, llgmi_params *params, const llgmi_material *mat,
                          int field_mask, double max_err, double headroom, double dt_min,
                          double dt_max, double *eps, int *accepted, double *dt_used) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m) return fail(LLGMI_ERR_INVALID, "m is NULL");
    std::string e = check_step_args(params, eps, accepted, dt_used, dt_min, dt_max);
    if (e.empty()) e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            eng.ensure_io();
            eng.upload(m, eng.d_m);
            eng.rk23(eng.d_m, *params, *mat, field_mask, max_err, headroom, dt_min, dt_max, *eps,
                     *accepted, *dt_used);
            if (*accepted) eng.download(eng.d_m, m);
        });
    });
}

int llgmi_dev_rk23_relax_step(llgmi_ctx *ctx, void *m_dev, llgmi_params *params,
                              const llgmi_material *mat, int field_mask, double max_err,
                              double headroom, double dt_min, double dt_max, double *eps,
                              int *accepted, double *dt_used) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    std::string e = check_step_args(params, eps, accepted, dt_used, dt_min, dt_max);
    if (e.empty()) e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.rk23((T *)m_dev, *params, *mat, field_mask, max_err, headroom, dt_min, dt_max, *eps,
                     *accepted, *dt_used);
        });
    });
}

int llgmi_rk23_invalidate(llgmi_ctx *ctx) {
    return guarded(ctx, [&] { with_engine(ctx, [&](auto &eng) { eng.rk23_invalidate(); }); });
}

int llgmi_rk23_last_b_eff(llgmi_ctx *ctx, double *b_out, int *valid) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!b_out || !valid) return fail(LLGMI_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            *valid = eng.r23_b4_valid ? 1 : 0;
            if (eng.r23_b4_valid) {
                eng.ensure_io();
                eng.download(eng.d_r23b4, b_out);
            }
        });
    });
}

int llgmi_compute_energy_geom(llgmi_ctx *ctx, const double *m, const llgmi_material *mat,
                              const double b_ext[3], const uint8_t *geom_mask, llgmi_energy *out) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m || !b_ext || !out) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mat(mat);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            eng.ensure_io();
            eng.upload(m, eng.d_m);
            const uint8_t *dg = eng.upload_geom(geom_mask);
            double r[5];
            eng.energy(eng.d_m, *mat, b_ext, dg, r);
            *out = llgmi_energy{r[0], r[1], r[2], r[3], r[4]};
        });
    });
}

int llgmi_dev_compute_energy_geom(llgmi_ctx *ctx, const void *m_dev, const llgmi_material *mat,
                                  const double b_ext[3], const uint8_t *geom_mask_dev,
                                  llgmi_energy *out) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!b_ext || !out) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_mat(mat);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (e.empty() && geom_mask_dev) e = check_dev(geom_mask_dev, n, "geom_mask");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            double r[5];
            eng.energy((const T *)m_dev, *mat, b_ext, geom_mask_dev, r);
            *out = llgmi_energy{r[0], r[1], r[2], r[3], r[4]};
        });
    });
}

int llgmi_relax_with_report(llgmi_ctx *ctx, double *m, llgmi_params *params, const llgmi_material *mat,
                            int field_mask, llgmi_relax_settings *settings, llgmi_relax_report *report) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!m || !params || !report) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_relax_settings(settings);
    if (e.empty()) e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            eng.ensure_io();
            eng.upload(m, eng.d_m);
            eng.relax(eng.d_m, *params, *mat, field_mask, *settings, *report);
            eng.download(eng.d_m, m);
        });
    });
}

int llgmi_dev_relax_with_report(llgmi_ctx *ctx, void *m_dev, llgmi_params *params,
                                const llgmi_material *mat, int field_mask,
                                llgmi_relax_settings *settings, llgmi_relax_report *report) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!params || !report) return fail(LLGMI_ERR_INVALID, "NULL argument");
    std::string e = check_relax_settings(settings);
    if (e.empty()) e = check_mat(mat);
    if (e.empty()) e = check_mask(field_mask);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.relax((T *)m_dev, *params, *mat, field_mask, *settings, *report);
        });
    });
}

// ---- on-disk formats --------------------------------------------------------------
namespace {
std::string check_path(const char *path) {
    if (!path || !*path) return "path is NULL or empty";
    return "";
}
int copy_string(const std::string &s, char *buf, size_t len) {
    if (!buf) return fail(LLGMI_ERR_INVALID, "buffer is NULL");
    if (s.size() + 1 > len) return fail(LLGMI_ERR_INVALID, "buffer too small (" + std::to_string(s.size() + 1) + " bytes needed)");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return LLGMI_OK;
}
extern "C++" template <typename F>
int guarded_host(F &&f) {
    try {
        f();
        return LLGMI_OK;
    } catch (const Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(LLGMI_ERR_ALLOC, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(LLGMI_ERR_INVALID, e.what());
    }
}
}  // namespace

int llgmi_write_ovf2_text(const char *path, const llgmi_grid *grid, const double *m, const llgmi_ovf_meta *meta) {
    std::string e = check_path(path);
    if (e.empty()) e = check_grid(grid);
    if (e.empty() && !m) e = "m is NULL";
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded_host([&] { write_ovf2_text(path, *grid, m, meta); });
}

int llgmi_write_ovf2_binary4(const char *path, const llgmi_grid *grid, const double *m,
                             const llgmi_ovf_meta *meta) {
    std::string e = check_path(path);
    if (e.empty()) e = check_grid(grid);
    if (e.empty() && !m) e = "m is NULL";
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded_host([&] { write_ovf2_binary4(path, *grid, m, meta); });
}

int llgmi_dev_write_ovf2(llgmi_ctx *ctx, const void *m_dev, const char *path, const llgmi_ovf_meta *meta,
                         int binary4) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    std::string e = check_path(path);
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    if (e.empty()) e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.ensure_io();
            std::vector<double> h((size_t)(3 * eng.n));
            eng.download((const T *)m_dev, h.data());
            if (binary4) write_ovf2_binary4(path, ctx->grid, h.data(), meta);
            else write_ovf2_text(path, ctx->grid, h.data(), meta);
        });
    });
}

int llgmi_ovf_total_sim_time_line(double t_s, char *buf, size_t len) {
    return copy_string(ovf_total_sim_time_line(t_s), buf, len);
}

int llgmi_demag_cache_path(const llgmi_grid *grid, uint64_t pbc_x, uint64_t pbc_y, char *buf, size_t len) {
    std::string e = check_grid(grid);
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return copy_string(kernel_cache_path(*grid, (int64_t)pbc_x, (int64_t)pbc_y), buf, len);
}

int llgmi_demag_cache_save(llgmi_ctx *ctx, uint64_t pbc_x, uint64_t pbc_y, const char *path) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            auto &op = eng.fft_for((int64_t)pbc_x, (int64_t)pbc_y);
            const std::string p = path ? std::string(path) : kernel_cache_path(ctx->grid, (int64_t)pbc_x, (int64_t)pbc_y);
            write_kernel_cache(p, ctx->grid, (int64_t)pbc_x, (int64_t)pbc_y, op.base_host);
        });
    });
}

int llgmi_demag_cache_load(llgmi_ctx *ctx, uint64_t pbc_x, uint64_t pbc_y, const char *path, int *loaded) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!loaded) return fail(LLGMI_ERR_INVALID, "loaded is NULL");
    if (pbc_x > (uint64_t)(1 << 20) || pbc_y > (uint64_t)(1 << 20))
        return fail(LLGMI_ERR_UNSUPPORTED, "periodic image count beyond 2^20");
    *loaded = 0;
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            const std::string p = path ? std::string(path) : kernel_cache_path(ctx->grid, (int64_t)pbc_x, (int64_t)pbc_y);
            std::vector<double> base;
            if (!read_kernel_cache(p, ctx->grid, (int64_t)pbc_x, (int64_t)pbc_y, base)) return;
            auto &op = (pbc_x == 0 && pbc_y == 0) ? eng.fft : eng.fft_pbc;
            eng.rebuild_fft(op, (int64_t)pbc_x, (int64_t)pbc_y, &base);
            *loaded = 1;
        });
    });
}

int llgmi_profile_fft_passes(llgmi_ctx *ctx, const void *m_dev, double ms, int reps, double *avg_us) {
    if (!ctx) return fail(LLGMI_ERR_INVALID, "context is NULL");
    if (!avg_us) return fail(LLGMI_ERR_INVALID, "avg_us is NULL");
    if (reps < 1 || reps > 100000) return fail(LLGMI_ERR_INVALID, "reps must be in [1, 100000]");
    const size_t n = (size_t)(ctx->grid.nx * ctx->grid.ny);
    std::string e = check_dev(m_dev, 3 * n * elem_size(ctx), "m");
    if (!e.empty()) return fail(LLGMI_ERR_INVALID, e);
    return guarded(ctx, [&] {
        with_engine(ctx, [&](auto &eng) {
            using T = std::remove_pointer_t<decltype(eng.d_m)>;
            eng.ensure_fft();
            eng.alloc(eng.d_bd, 3 * eng.n);
            auto &op = eng.fft;
            hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
            auto destroy = [&] {
                for (auto &x : ev)
                    if (x) (void)hipEventDestroy(x);
            };
            try {
                for (auto &x : ev) LLGMI_HIP(hipEventCreate(&x));
                LLGMI_HIP(hipEventRecord(ev[0], eng.stream));
                for (int i = 0; i < reps; ++i) op.launch_row_fwd((const T *)m_dev, T(ms), eng.stream, op.rq_row);
                LLGMI_HIP(hipEventRecord(ev[1], eng.stream));
                for (int i = 0; i < reps; ++i) op.launch_col(eng.stream, op.rq_col);
                LLGMI_HIP(hipEventRecord(ev[2], eng.stream));
                for (int i = 0; i < reps; ++i) op.launch_row_inv(eng.d_bd, eng.stream, op.rq_row_inv);
                LLGMI_HIP(hipEventRecord(ev[3], eng.stream));
                LLGMI_HIP(hipEventSynchronize(ev[3]));
                for (int k = 0; k < 3; ++k) {
                    float t = 0.f;
                    LLGMI_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
                    avg_us[k] = 1e3 * (double)t / reps;
                }
            } catch (...) {
                destroy();
                throw;
            }
            destroy();
        });
    });
}

int llgmi_profile_enable(llgmi_ctx *ctx, int class_mask) {
    return guarded(ctx, [&] { ctx->prof.mask = (unsigned)class_mask & ((1u << LLGMI_K_COUNT) - 1u); });
}

int llgmi_profile_reset(llgmi_ctx *ctx) {
    return guarded(ctx, [&] { ctx->prof.reset(); });
}

int llgmi_profile_read(llgmi_ctx *ctx, int cls, double *total_ms, int64_t *count) {
    if (cls < 0 || cls >= LLGMI_K_COUNT) return fail(LLGMI_ERR_INVALID, "unknown kernel class");
    if (!total_ms || !count) return fail(LLGMI_ERR_INVALID, "NULL argument");
    return guarded(ctx, [&] {
        ctx->prof.drain();
        *total_ms = ctx->prof.total_ms[cls];
        *count = ctx->prof.count[cls];
    });
}

}  // extern "C"
