// llg_sim.hpp -- C++ host API over libllgmi.so with the reference's names.
//
// Mirrors llg-sim's Rust API for the uniform-grid hot path, one-to-one:
//   Grid2D (grid.rs:5-31), VectorField2D (vector_field.rs:24-52),
//   Material / LLGParams / DemagMethod / MU0 / GAMMA_E_RAD_PER_S_T
//   (params.rs:15-16, 67-144), FieldMask (effective_field/mod.rs:37-45),
//   effective_field::build_h_eff_masked[_geom] / build_h_eff (mod.rs:48-112),
//   effective_field::demag::{add,compute}_demag_field[_pbc] (demag.rs:47-156),
//   llg::RK45Scratch (llg.rs:126-189),
//   llg::step_llg_rk45_recompute_field[_masked]_adaptive (llg.rs:795-1050),
//   llg::RK23Scratch / step_llg_rk23_recompute_field_masked_relax_adaptive
//   (llg.rs:196-436), energy::compute_energy[_geom] / compute_total_energy
//   (energy.rs:21-305), relax::RelaxSettings / relax_with_report / relax /
//   mumax_like_relax_settings (relax.rs:32-808), ovf::OvfMeta /
//   write_ovf2_rectangular_{text,binary4} / OvfSeries (ovf.rs:17-316).
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
#include <utility>
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

// RK23Scratch::new(grid) (llg.rs:223-247): relax-mode stage storage, FSAL and
// last-field state live on the GPU.
class RK23Scratch {
public:
    explicit RK23Scratch(Grid2D grid, int precision = LLGMI_F64) : grid_(grid), ctx_(grid, precision) {}
    llgmi_ctx *ctx() const { return ctx_.get(); }
    const Grid2D &grid() const { return grid_; }
    // last_b_eff (llg.rs:252-258): the field of the accepted state, if held.
    std::optional<VectorField2D> last_b_eff() const {
        VectorField2D b(grid_);
        int valid = 0;
        check(llgmi_rk23_last_b_eff(ctx_.get(), b.ptr(), &valid), "last_b_eff");
        if (!valid) return std::nullopt;
        return b;
    }
    // invalidate_last_b_eff (llg.rs:264-268)
    void invalidate_last_b_eff() { check(llgmi_rk23_invalidate(ctx_.get()), "invalidate_last_b_eff"); }

private:
    Grid2D grid_;
    Context ctx_;
};

// step_llg_rk23_recompute_field_masked_relax_adaptive (llg.rs:278-436).
inline std::tuple<double, bool, double> step_llg_rk23_recompute_field_masked_relax_adaptive(
    VectorField2D &m, LLGParams &params, const Material &material, RK23Scratch &scratch,
    FieldMask mask, double max_err, double headroom, double dt_min, double dt_max) {
    if (m.data.size() != scratch.grid().n_cells()) throw std::runtime_error("grid mismatch");
    llgmi_params p = params.c();
    llgmi_material cm = material.c();
    double eps = 0, dt_used = 0;
    int accepted = 0;
    check(llgmi_rk23_relax_step(scratch.ctx(), m.ptr(), &p, &cm, (int)mask, max_err, headroom, dt_min,
                                dt_max, &eps, &accepted, &dt_used),
          "step_llg_rk23_recompute_field_masked_relax_adaptive");
    params.dt = p.dt;
    return {eps, accepted != 0, dt_used};
}

}  // namespace llg

// energy.rs:21-305
namespace energy {

struct EnergyBreakdown {
    double exchange = 0, anisotropy = 0, zeeman = 0, dmi = 0, demag = 0;
    double total() const { return exchange + anisotropy + zeeman + dmi + demag; }
};

inline EnergyBreakdown compute_energy_geom(const Grid2D &grid, const VectorField2D &m, const Material &material,
                                           const Vec3 &b_ext, const std::vector<uint8_t> *geom_mask) {
    if (m.data.size() != grid.n_cells()) throw std::runtime_error("field length must match grid.n_cells()");
    if (geom_mask && geom_mask->size() != grid.n_cells())
        throw std::runtime_error("mask length must match grid.n_cells()");
    auto &ctx = detail::shared_context(grid, detail::default_precision());
    llgmi_material cm = material.c();
    llgmi_energy e{};
    check(llgmi_compute_energy_geom(ctx.get(), m.ptr(), &cm, b_ext.data(),
                                    geom_mask ? geom_mask->data() : nullptr, &e),
          "compute_energy_geom");
    return EnergyBreakdown{e.exchange, e.anisotropy, e.zeeman, e.dmi, e.demag};
}
inline EnergyBreakdown compute_energy(const Grid2D &grid, const VectorField2D &m, const Material &material,
                                      const Vec3 &b_ext) {
    return compute_energy_geom(grid, m, material, b_ext, nullptr);
}
inline double compute_total_energy(const Grid2D &grid, const VectorField2D &m, const Material &material,
                                   const Vec3 &b_ext) {
    return compute_energy(grid, m, material, b_ext).total();
}

}  // namespace energy

// relax.rs:32-808
namespace relax {

enum class TorqueMetric { Max = LLGMI_TORQUE_MAX, Mean = LLGMI_TORQUE_MEAN, Rms = LLGMI_TORQUE_RMS,
                          MeanSq = LLGMI_TORQUE_MEANSQ };
enum class RelaxStopReason { MaxAcceptedSteps = LLGMI_STOP_MAX_ACCEPTED_STEPS,
                             Phase2Disabled = LLGMI_STOP_PHASE2_DISABLED,
                             TightenFloorReached = LLGMI_STOP_TIGHTEN_FLOOR_REACHED,
                             TorqueGateNotSatisfied = LLGMI_STOP_TORQUE_GATE_NOT_SATISFIED };
enum class RelaxStageStop { BelowThreshold = LLGMI_STAGE_BELOW_THRESHOLD, Plateau = LLGMI_STAGE_PLATEAU,
                            MaxAcceptedSteps = LLGMI_STAGE_MAX_ACCEPTED_STEPS };

struct RelaxReport {
    size_t accepted_steps = 0, rejected_steps = 0, torque_checks = 0, torque_field_rebuilds = 0;
    RelaxStopReason stop_reason = RelaxStopReason::MaxAcceptedSteps;
    std::optional<RelaxStageStop> last_stage_stop;
    std::optional<double> final_torque, final_torque_max;
    double final_max_err = 0, final_dt = 0;
};

struct RelaxSettings {
    double max_err = 1e-5, headroom = 0.8, dt_min = 1e-18, dt_max = 1e-11;
    bool phase1_enabled = true, phase2_enabled = true;
    size_t energy_stride = 3;
    double rel_energy_tol = 1e-12;
    TorqueMetric torque_metric = TorqueMetric::Max;
    std::optional<double> torque_threshold = 1e-4;
    size_t torque_check_stride = 1;
    double torque_plateau_rel = 1e-3, torque_plateau_abs = 0.0;
    size_t torque_plateau_checks = 0, torque_plateau_min_checks = 5;
    double tighten_factor = 0.70710678118654752440, tighten_floor = 1e-9;
    size_t max_accepted_steps = 2000000;
    std::optional<double> final_torque_gate, final_torque_gate_max;
    size_t gate_max_extra_accepted_steps = 0, gate_plateau_fails = 0;

    llgmi_relax_settings c() const {
        llgmi_relax_settings s{};
        s.max_err = max_err; s.headroom = headroom; s.dt_min = dt_min; s.dt_max = dt_max;
        s.phase1_enabled = phase1_enabled; s.phase2_enabled = phase2_enabled;
        s.energy_stride = (int64_t)energy_stride; s.rel_energy_tol = rel_energy_tol;
        s.torque_metric = (int32_t)torque_metric;
        s.has_torque_threshold = torque_threshold.has_value();
        s.torque_threshold = torque_threshold.value_or(0.0);
        s.torque_check_stride = (int64_t)torque_check_stride;
        s.torque_plateau_rel = torque_plateau_rel; s.torque_plateau_abs = torque_plateau_abs;
        s.torque_plateau_checks = (int64_t)torque_plateau_checks;
        s.torque_plateau_min_checks = (int64_t)torque_plateau_min_checks;
        s.tighten_factor = tighten_factor; s.tighten_floor = tighten_floor;
        s.max_accepted_steps = (int64_t)max_accepted_steps;
        s.has_final_torque_gate = final_torque_gate.has_value();
        s.has_final_torque_gate_max = final_torque_gate_max.has_value();
        s.final_torque_gate = final_torque_gate.value_or(0.0);
        s.final_torque_gate_max = final_torque_gate_max.value_or(0.0);
        s.gate_max_extra_accepted_steps = (int64_t)gate_max_extra_accepted_steps;
        s.gate_plateau_fails = (int64_t)gate_plateau_fails;
        return s;
    }
};

// mumax_like_relax_settings (relax.rs:789-808)
inline RelaxSettings mumax_like_relax_settings() {
    RelaxSettings s;
    s.rel_energy_tol = 0.0;
    s.torque_metric = TorqueMetric::MeanSq;
    s.torque_threshold.reset();
    s.torque_plateau_checks = 1;
    s.torque_plateau_min_checks = 1;
    s.torque_plateau_rel = 0.0;
    s.torque_plateau_abs = 0.0;
    return s;
}

// relax_with_report (relax.rs:334-774): settings.max_err is written back.
inline RelaxReport relax_with_report(const Grid2D &grid, VectorField2D &m, LLGParams &params,
                                     const Material &material, llg::RK23Scratch &scratch, FieldMask mask,
                                     RelaxSettings &settings) {
    if (m.data.size() != grid.n_cells() || scratch.grid().n_cells() != grid.n_cells())
        throw std::runtime_error("grid mismatch");
    llgmi_params p = params.c();
    llgmi_material cm = material.c();
    llgmi_relax_settings cs = settings.c();
    llgmi_relax_report r{};
    check(llgmi_relax_with_report(scratch.ctx(), m.ptr(), &p, &cm, (int)mask, &cs, &r), "relax_with_report");
    params.dt = p.dt;
    settings.max_err = cs.max_err;
    RelaxReport out;
    out.accepted_steps = (size_t)r.accepted_steps;
    out.rejected_steps = (size_t)r.rejected_steps;
    out.torque_checks = (size_t)r.torque_checks;
    out.torque_field_rebuilds = (size_t)r.torque_field_rebuilds;
    out.stop_reason = (RelaxStopReason)r.stop_reason;
    if (r.last_stage_stop != LLGMI_STAGE_NONE) out.last_stage_stop = (RelaxStageStop)r.last_stage_stop;
    if (r.has_final_torque) out.final_torque = r.final_torque;
    if (r.has_final_torque_max) out.final_torque_max = r.final_torque_max;
    out.final_max_err = r.final_max_err;
    out.final_dt = r.final_dt;
    return out;
}

inline void relax(const Grid2D &grid, VectorField2D &m, LLGParams &params, const Material &material,
                  llg::RK23Scratch &scratch, FieldMask mask, RelaxSettings &settings) {
    (void)relax_with_report(grid, m, params, material, scratch, mask, settings);
}

}  // namespace relax

// ovf.rs:17-316
namespace ovf {

struct OvfMeta {
    std::string title;
    std::vector<std::string> desc_lines;
    std::array<std::string, 3> valuelabels, valueunits;

    static OvfMeta magnetization() {
        OvfMeta m;
        m.title = "m";
        m.valuelabels = {"m_x", "m_y", "m_z"};
        m.valueunits = {"1", "1", "1"};
        return m;
    }
    OvfMeta with_total_sim_time(double t_s) && {
        char buf[128];
        check(llgmi_ovf_total_sim_time_line(t_s, buf, sizeof buf), "with_total_sim_time");
        desc_lines.emplace_back(buf);
        return std::move(*this);
    }
    void push_desc_line(std::string s) { desc_lines.push_back(std::move(s)); }
};

namespace detail {
template <typename F>
inline void with_c_meta(const OvfMeta &meta, F &&f) {
    std::vector<const char *> lines;
    for (auto &d : meta.desc_lines) lines.push_back(d.c_str());
    llgmi_ovf_meta c{};
    c.title = meta.title.c_str();
    c.desc_lines = lines.empty() ? nullptr : lines.data();
    c.n_desc = (int32_t)lines.size();
    for (int q = 0; q < 3; ++q) {
        c.valuelabels[q] = meta.valuelabels[q].c_str();
        c.valueunits[q] = meta.valueunits[q].c_str();
    }
    f(c);
}
}  // namespace detail

inline void write_ovf2_rectangular_text(const std::string &path, const Grid2D &grid, const VectorField2D &m,
                                        const OvfMeta &meta) {
    llgmi_grid g{(int64_t)grid.nx, (int64_t)grid.ny, grid.dx, grid.dy, grid.dz};
    detail::with_c_meta(meta, [&](const llgmi_ovf_meta &c) {
        check(llgmi_write_ovf2_text(path.c_str(), &g, m.ptr(), &c), "write_ovf2_rectangular_text");
    });
}

inline void write_ovf2_rectangular_binary4(const std::string &path, const Grid2D &grid, const VectorField2D &m,
                                           const OvfMeta &meta) {
    llgmi_grid g{(int64_t)grid.nx, (int64_t)grid.ny, grid.dx, grid.dy, grid.dz};
    detail::with_c_meta(meta, [&](const llgmi_ovf_meta &c) {
        check(llgmi_write_ovf2_binary4(path.c_str(), &g, m.ptr(), &c), "write_ovf2_rectangular_binary4");
    });
}

// OvfSeries (ovf.rs:264-316): binary4 dumps every `every` seconds.
class OvfSeries {
public:
    std::string dir;
    double every;
    OvfSeries(std::string dir_, double every_s) : dir(std::move(dir_)), every(every_s > 0 ? every_s : 0.0) {}
    template <typename MakePath>
    void dump_initial(const Grid2D &grid, const VectorField2D &m, MakePath make_path, const OvfMeta &meta) {
        next_dump_ = 0.0;
        write_ovf2_rectangular_binary4(make_path(frame_, 0.0), grid, m, meta);
        ++frame_;
        next_dump_ = every;
    }
    template <typename MakePath, typename MetaBuilder>
    void maybe_dump_due(const Grid2D &grid, const VectorField2D &m, double t_approx, MakePath make_path,
                        MetaBuilder meta_builder) {
        if (every <= 0.0) return;
        const double eps = 1e-15;
        while (t_approx + eps >= next_dump_) {
            const double t = next_dump_;
            write_ovf2_rectangular_binary4(make_path(frame_, t), grid, m, meta_builder(t));
            ++frame_;
            next_dump_ += every;
        }
    }
    uint64_t frame() const { return frame_; }

private:
    double next_dump_ = 0.0;
    uint64_t frame_ = 0;
};

}  // namespace ovf
}  // namespace llg_sim
