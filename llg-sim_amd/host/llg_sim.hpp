// llg_sim.hpp -- C++ host API over libllgmi.so with the reference's names.
//
// Mirrors llg-sim's Rust API for the uniform-grid hot path, one-to-one:
//   Grid2D (grid.rs:5-31), VectorField2D (vector_field.rs:24-52),
//   Material / LLGParams / DemagMethod / MU0 / GAMMA_E_RAD_PER_S_T
//   (params.rs:15-16, 67-144), FieldMask (effective_field/mod.rs:37-45),
//   effective_field::build_h_eff_masked[_geom] / build_h_eff (mod.rs:48-112),
//   effective_field::demag::{add,compute}_demag_field[_pbc] (demag.rs:47-156),
//   llg::RK45Scratch (llg.rs:126-189),
//   llg::step_llg_rk45_recompute_field[_masked]_adaptive (llg.rs:795-1050).
//
// Field builders use a process-wide cache of GPU contexts keyed by grid and
// precision (the reference keys its demag operator cache the same way,
// demag_fft_uniform.rs:44-108); RK45Scratch owns a context of its own, which
// carries the FSAL state.  Failures throw std::runtime_error (the reference
// panics); data stay in the reference's AoS f64 layout.
#pragma once

#include "../../include/llgmi.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace llg_sim {

constexpr double PI = 3.14159265358979323846;
constexpr double MU0 = 4.0 * PI * 1e-7;
constexpr double GAMMA_E_RAD_PER_S_T = 1.76085963023e11;

using Vec3 = std::array<double, 3>;

struct Grid2D {
    size_t nx = 0, ny = 0;
    double dx = 0, dy = 0, dz = 0;
    Grid2D() = default;
    Grid2D(size_t nx_, size_t ny_, double dx_, double dy_, double dz_)
        : nx(nx_), ny(ny_), dx(dx_), dy(dy_), dz(dz_) {}
    size_t n_cells() const { return nx * ny; }
    double cell_volume() const { return dx * dy * dz; }
    size_t idx(size_t i, size_t j) const { return j * nx + i; }
};

struct VectorField2D {
    Grid2D grid;
    std::vector<Vec3> data;
    explicit VectorField2D(Grid2D g) : grid(g), data(g.n_cells(), Vec3{0.0, 0.0, 1.0}) {}
    void set_uniform(double x, double y, double z) {
        for (auto &c : data) c = {x, y, z};
    }
    size_t idx(size_t i, size_t j) const { return grid.idx(i, j); }
    const double *ptr() const { return data.empty() ? nullptr : data[0].data(); }
    double *ptr() { return data.empty() ? nullptr : data[0].data(); }
};

enum class DemagMethod { FftUniform = LLGMI_DEMAG_FFT, PoissonMG = LLGMI_DEMAG_MG, PoissonDst = LLGMI_DEMAG_DST };

struct Material {
    double ms = 0, a_ex = 0, k_u = 0;
    Vec3 easy_axis{0.0, 0.0, 1.0};
    std::optional<double> dmi;
    bool demag = false;
    DemagMethod demag_method = DemagMethod::FftUniform;

    llgmi_material c() const {
        llgmi_material m{};
        m.ms = ms; m.a_ex = a_ex; m.k_u = k_u;
        for (int q = 0; q < 3; ++q) m.easy_axis[q] = easy_axis[q];
        m.has_dmi = dmi.has_value() ? 1 : 0;
        m.dmi = dmi.value_or(0.0);
        m.demag = demag ? 1 : 0;
        m.demag_method = (int32_t)demag_method;
        return m;
    }
};

struct LLGParams {
    double gamma = GAMMA_E_RAD_PER_S_T, alpha = 0, dt = 0;
    Vec3 b_ext{0.0, 0.0, 0.0};

    llgmi_params c() const {
        llgmi_params p{};
        p.gamma = gamma; p.alpha = alpha; p.dt = dt;
        for (int q = 0; q < 3; ++q) p.b_ext[q] = b_ext[q];
        return p;
    }
};

enum class FieldMask { ExchAnis = LLGMI_MASK_EXCH_ANIS, ExchAnisDmi = LLGMI_MASK_EXCH_ANIS_DMI, Full = LLGMI_MASK_FULL };

inline void check(int rc, const char *what) {
    if (rc != LLGMI_OK) throw std::runtime_error(std::string(what) + ": " + llgmi_last_error());
}

// Owning handle of one llgmi context.
class Context {
public:
    Context(const Grid2D &g, int precision = LLGMI_F64, int device = 0) {
        llgmi_grid cg{(int64_t)g.nx, (int64_t)g.ny, g.dx, g.dy, g.dz};
        check(llgmi_ctx_create(&cg, precision, device, &h_), "llgmi_ctx_create");
    }
    ~Context() { llgmi_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    llgmi_ctx *get() const { return h_; }

private:
    llgmi_ctx *h_ = nullptr;
};

namespace detail {
inline Context &shared_context(const Grid2D &g, int precision) {
    struct Key {
        size_t nx, ny;
        double dx, dy, dz;
        int prec;
        bool operator<(const Key &o) const {
            return std::tie(nx, ny, dx, dy, dz, prec) < std::tie(o.nx, o.ny, o.dx, o.dy, o.dz, o.prec);
        }
    };
    static std::mutex mu;
    static std::map<Key, std::unique_ptr<Context>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto &slot = cache[Key{g.nx, g.ny, g.dx, g.dy, g.dz, precision}];
    if (!slot) slot.reset(new Context(g, precision));
    return *slot;
}
inline int &default_precision() {
    static int p = LLGMI_F64;
    return p;
}
}  // namespace detail

// Precision used by the free-function field builders (FP64 by default).
inline void set_default_precision(int precision) { detail::default_precision() = precision; }

namespace effective_field {

inline void build_h_eff_masked_geom(const Grid2D &grid, const VectorField2D &m, VectorField2D &b_eff,
                                    const LLGParams &sim, const Material &mat, FieldMask mask,
                                    const std::vector<uint8_t> *geom_mask) {
    if (m.data.size() != grid.n_cells() || b_eff.data.size() != grid.n_cells())
        throw std::runtime_error("field length must match grid.n_cells()");
    if (geom_mask && geom_mask->size() != grid.n_cells())
        throw std::runtime_error("mask length must match grid.n_cells()");
    auto &ctx = detail::shared_context(grid, detail::default_precision());
    llgmi_params p = sim.c();
    llgmi_material cm = mat.c();
    check(llgmi_build_h_eff_masked_geom(ctx.get(), m.ptr(), b_eff.ptr(), &p, &cm, (int)mask,
                                        geom_mask ? geom_mask->data() : nullptr),
          "build_h_eff_masked_geom");
}

inline void build_h_eff_masked(const Grid2D &grid, const VectorField2D &m, VectorField2D &b_eff,
                               const LLGParams &sim, const Material &mat, FieldMask mask) {
    build_h_eff_masked_geom(grid, m, b_eff, sim, mat, mask, nullptr);
}

inline void build_h_eff(const Grid2D &grid, const VectorField2D &m, VectorField2D &b_eff,
                        const LLGParams &sim, const Material &mat) {
    build_h_eff_masked(grid, m, b_eff, sim, mat, FieldMask::Full);
}

namespace demag {
inline void add_demag_field_pbc(const Grid2D &grid, const VectorField2D &m, VectorField2D &b_eff,
                                const Material &mat, size_t pbc_x, size_t pbc_y) {
    auto &ctx = detail::shared_context(grid, detail::default_precision());
    llgmi_material cm = mat.c();
    check(llgmi_add_demag_field_pbc(ctx.get(), m.ptr(), b_eff.ptr(), &cm, pbc_x, pbc_y),
          "add_demag_field_pbc");
}
inline void add_demag_field(const Grid2D &grid, const VectorField2D &m, VectorField2D &b_eff,
                            const Material &mat) {
    add_demag_field_pbc(grid, m, b_eff, mat, 0, 0);
}
inline void compute_demag_field_pbc(const Grid2D &grid, const VectorField2D &m, VectorField2D &out,
                                    const Material &mat, size_t pbc_x, size_t pbc_y) {
    out.set_uniform(0.0, 0.0, 0.0);
    add_demag_field_pbc(grid, m, out, mat, pbc_x, pbc_y);
}
inline void compute_demag_field(const Grid2D &grid, const VectorField2D &m, VectorField2D &out,
                                const Material &mat) {
    compute_demag_field_pbc(grid, m, out, mat, 0, 0);
}
}  // namespace demag

}  // namespace effective_field

namespace llg {

// RK45Scratch::new(grid) (llg.rs:159): stage storage and FSAL state live on the GPU.
class RK45Scratch {
public:
    explicit RK45Scratch(Grid2D grid, int precision = LLGMI_F64) : grid_(grid), ctx_(grid, precision) {}
    llgmi_ctx *ctx() const { return ctx_.get(); }
    const Grid2D &grid() const { return grid_; }

private:
    Grid2D grid_;
    Context ctx_;
};

// Returns (eps, accepted, dt_used); params.dt is replaced by the next step size.
inline std::tuple<double, bool, double> step_llg_rk45_recompute_field_masked_adaptive(
    VectorField2D &m, LLGParams &params, const Material &material, RK45Scratch &scratch,
    FieldMask mask, double max_err, double headroom, double dt_min, double dt_max) {
    if (m.data.size() != scratch.grid().n_cells()) throw std::runtime_error("grid mismatch");
    llgmi_params p = params.c();
    llgmi_material cm = material.c();
    double eps = 0, dt_used = 0;
    int accepted = 0;
    check(llgmi_rk45_step(scratch.ctx(), m.ptr(), &p, &cm, (int)mask, max_err, headroom, dt_min,
                          dt_max, &eps, &accepted, &dt_used),
          "step_llg_rk45_recompute_field_masked_adaptive");
    params.dt = p.dt;
    return {eps, accepted != 0, dt_used};
}

inline std::tuple<double, bool, double> step_llg_rk45_recompute_field_adaptive(
    VectorField2D &m, LLGParams &params, const Material &material, RK45Scratch &scratch,
    double max_err, double headroom, double dt_min, double dt_max) {
    return step_llg_rk45_recompute_field_masked_adaptive(m, params, material, scratch, FieldMask::Full,
                                                          max_err, headroom, dt_min, dt_max);
}

}  // namespace llg
}  // namespace llg_sim
