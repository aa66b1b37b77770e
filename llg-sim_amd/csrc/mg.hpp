// Geometric-multigrid Poisson demag (the second demag path of llg-sim,
// demag_poisson_mg.rs + mg_kernels.rs + mg_treecode.rs), FP64 on the GPU.
//
// The padded box, level hierarchy, stencils and solve loop follow the
// reference; the box arrays of every level live in HBM (index
// (k*ny + j)*nx + i), the smoothers, transfers, residuals, RHS, gradient and
// boundary-value evaluation are HIP kernels.  The treecode octree is built on
// the host from the compacted charge list only when the set of charged cells
// changes; its moments and the boundary evaluation run on the GPU every solve.
#pragma once

#include "internal.hpp"
#include "../../include/llgmi.h"

#include <array>
#include <map>
#include <string>
#include <vector>

namespace llgmi {

// Stencil3D (demag_poisson_mg.rs:810-816)
struct MGStencil {
    double center = 0.0, diag = 0.0;
    std::vector<std::array<int, 3>> offs;
    std::vector<double> coeffs;
};

MGStencil mg_stencil_from_kind(int kind, double dx, double dy, double dz, double alpha);
MGStencil mg_galerkin(const MGStencil &fine, int rx, int ry, int rz, bool trilinear);
void mg_config_default(llgmi_mg_config &c);
void mg_config_from_env(llgmi_mg_config &c);
void mg_hybrid_config_default(llgmi_mg_hybrid_config &h);
void mg_hybrid_config_from_env(llgmi_mg_hybrid_config &h);
void mg_sanitize(llgmi_mg_config &c);
void mg_validate(const llgmi_mg_config &c);  // throws Error(1) on out-of-range fields

class MGSolver {
public:
    MGSolver(const llgmi_grid &g, const llgmi_mg_config &cfg, hipStream_t stream, Profiler *prof);
    ~MGSolver();
    MGSolver(const MGSolver &) = delete;
    MGSolver &operator=(const MGSolver &) = delete;

    bool same_structure(const llgmi_grid &g, const llgmi_mg_config &cfg) const;
    void apply_cfg(const llgmi_mg_config &cfg);
    // B_demag = -mu0 grad(phi) on the film cells with |m|^2 > 1e-30, 0 elsewhere
    // (overwrites bd); m and bd are SoA in precision T.  sigma > 0 screens the
    // rhs with a Gaussian (hybrid); masked = false fills every film cell.
    template <typename T>
    void demag_field(const T *m, double ms, T *bd, double sigma = 0.0, bool masked = true);
    const llgmi_mg_stats &stats() const { return stats_; }
    const llgmi_mg_config &config() const { return cfg_; }
    // finest box (px, py, pz, offx, offy, offz)
    void box(int64_t d[6]) const {
        d[0] = px_; d[1] = py_; d[2] = pz_; d[3] = offx_; d[4] = offy_; d[5] = offz_;
    }
    // device copy of the finest phi out of (into = false) or into the solver
    void phi_copy(double *dev, bool into);

    struct Level {
        int nx = 0, ny = 0, nz = 0;
        double dx = 0, dy = 0, dz = 0;
        MGStencil st;
        double *phi[2] = {nullptr, nullptr};  // [cur] is the level's phi; [1-cur] Jacobi target
        int cur = 0;
        double *rhs = nullptr;
        void *d_st = nullptr;                 // device copy of the stencil
        bool seven = false;                   // stencil is the 7-point layout (fast residual path)
    };
    struct Tree;

private:
    void release();
    void build_rhs(const void *m, bool f32, double ms);
    void update_bc();
    void treecode_bc();
    void dipole_bc();
    void stamp(int l);
    void smooth(int l, int64_t iters);
    void ensure_pair(Level &L);
    void v_cycle(int l);
    void v_cycle_graph();
    void solve();
    void screen_rhs(double sigma);
    double finest_rhs_max();
    template <typename T>
    void gradient(const T *m, T *bd);

    llgmi_grid g_{};
    llgmi_mg_config cfg_{};
    hipStream_t stream_;
    Profiler *prof_;
    int px_ = 0, py_ = 0, pz_ = 0, offx_ = 0, offy_ = 0, offz_ = 0;
    std::vector<Level> lev_;
    double *bc_ = nullptr;                     // finest boundary values (full box)
    double *d_partial_ = nullptr;              // reduction partials
    double *d_scalars_ = nullptr;              // [0] max, [1] nan flag, [2..5] dipole moments
    double *h_scalars_ = nullptr;              // pinned mirror
    Tree *tree_ = nullptr;
    double *d_screen_w_ = nullptr;             // Gaussian screening weights
    size_t screen_cap_ = 0;
    std::vector<double> screen_w_;
    llgmi_mg_stats stats_{};
    // V-cycles captured as HIP graphs, keyed by the smoothing settings and
    // the Jacobi buffer parity of every level at entry
    struct Graph {
        hipGraphExec_t exec = nullptr;
        std::vector<int> cur_after;
    };
    std::map<std::string, Graph> graphs_;
    bool use_graphs_ = true;   // LLGMI_MG_GRAPH=0: launch the V-cycle kernels directly
    bool rb_fused_ = true;     // LLGMI_MG_RB_FUSED=0: two half-sweep launches per sweep
    int rbz_ = 1 << 20;        // LLGMI_MG_RBZ: planes per z-chunk of the fused finest sweep
};

// Hybrid MG correction (DeltaKernel2D::add_correction, demag_poisson_mg.rs:410-583):
// b += dK * (ms m) over the (2r+1)^2 stencil dk = [xx | xy | yy | zz] (device,
// FP64); bscale > 0 scales cells near vacuum (scale_buf: nx*ny doubles).
template <typename T>
void launch_hybrid_correction(const double *dk, int r, const T *m, T *b, double ms, double bscale, int bmargin,
                              double *scale_buf, int nx, int ny, hipStream_t s);

}  // namespace llgmi
