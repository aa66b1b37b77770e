// The reference's own tests for this path, restated against the C++ host API
// (llg_sim.hpp over libllgmi.so).  Each test names the Rust test it mirrors.
// Exit status 0 when every test passes.
#include "../../llg-sim_amd/host/llg_sim.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <functional>

using namespace llg_sim;

static int g_fail = 0;
#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::printf("  FAIL %s:%d: %s  ", __FILE__, __LINE__, #cond); \
            std::printf(__VA_ARGS__);                                      \
            std::printf("\n");                                             \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

static Material demag_only(double ms) {
    Material m;
    m.ms = ms;
    m.demag = true;
    return m;
}

// demag_fft_uniform.rs:1261-1293
static void demag_single_cell_cube_is_close_to_minus_mu0_over_3() {
    Grid2D grid(1, 1, 1.0, 1.0, 1.0);
    VectorField2D m(grid);
    m.set_uniform(0.0, 0.0, 1.0);
    VectorField2D b(grid);
    b.set_uniform(0.0, 0.0, 0.0);
    effective_field::demag::add_demag_field(grid, m, b, demag_only(1.0));
    double bz = b.data[0][2], expected = -MU0 / 3.0;
    CHECK(std::fabs(bz - expected) < 1e-3, "bz=%g expected=%g", bz, expected);
}

// tests/validation.rs:478-552
static void demag_uniform_2x2_cube_cells_has_symmetry_and_reasonable_factor() {
    Grid2D grid(2, 2, 1.0, 1.0, 1.0);
    VectorField2D m(grid);
    m.set_uniform(0.0, 0.0, 1.0);
    VectorField2D b(grid);
    b.set_uniform(0.0, 0.0, 0.0);
    effective_field::demag::add_demag_field(grid, m, b, demag_only(1.0));
    double bz_min = b.data[0][2], bz_max = b.data[0][2];
    for (auto &v : b.data) {
        CHECK(std::fabs(v[0]) < 1e-10, "Bx=%g", v[0]);
        CHECK(std::fabs(v[1]) < 1e-10, "By=%g", v[1]);
        bz_min = std::fmin(bz_min, v[2]);
        bz_max = std::fmax(bz_max, v[2]);
    }
    CHECK(std::fabs(bz_max - bz_min) < 1e-10, "range %g..%g", bz_min, bz_max);
    CHECK(b.data[0][2] < 0.0, "bz=%g", b.data[0][2]);
    double nzz = -b.data[0][2] / MU0;
    CHECK(nzz > 0.05 && nzz < 1.05, "Nzz=%g", nzz);
}

// tests/validation.rs:558-615
static void demag_uniform_2x2_prints_nxx_nyy_nzz() {
    Grid2D grid(2, 2, 1.0, 1.0, 1.0);
    double n[3];
    for (int c = 0; c < 3; ++c) {
        VectorField2D m(grid);
        m.set_uniform(c == 0, c == 1, c == 2);
        VectorField2D b(grid);
        b.set_uniform(0.0, 0.0, 0.0);
        effective_field::demag::add_demag_field(grid, m, b, demag_only(1.0));
        n[c] = -b.data[0][c] / MU0;
    }
    std::printf("  2x2x1 implied demag factors: Nxx=%.6f, Nyy=%.6f, Nzz=%.6f\n", n[0], n[1], n[2]);
    for (int c = 0; c < 3; ++c) CHECK(n[c] > 0.0 && n[c] < 1.2, "N=%g", n[c]);
    double s = n[0] + n[1] + n[2];
    CHECK(s > 0.5 && s < 2.5, "sum=%g", s);
}

// tests/validation.rs:617-644 (E_demag = -1/2 sum M.B V, energy.rs)
static void demag_energy_is_nonnegative_for_uniform_state() {
    Grid2D grid(32, 32, 5e-9, 5e-9, 1e-9);
    VectorField2D m(grid);
    m.set_uniform(0.0, 0.0, 1.0);
    VectorField2D b(grid);
    effective_field::demag::compute_demag_field(grid, m, b, demag_only(1.0));
    double e = 0.0;
    for (size_t i = 0; i < m.data.size(); ++i)
        e += -0.5 * 1.0 * (m.data[i][0] * b.data[i][0] + m.data[i][1] * b.data[i][1] + m.data[i][2] * b.data[i][2]) *
             grid.cell_volume();
    CHECK(e >= 0.0, "E=%g", e);
}

// tests/validation.rs:395-472
static void dmi_field_flips_sign_with_d() {
    const size_t nx = 9;
    Grid2D grid(nx, 1, 1.0, 1.0, 1.0);
    VectorField2D m(grid);
    for (size_t i = 0; i < nx; ++i) {
        double mz = ((double)i - ((double)nx - 1.0) / 2.0) * 0.01;
        mz = std::fmin(std::fmax(mz, -0.2), 0.2);
        m.data[m.idx(i, 0)] = {std::sqrt(1.0 - mz * mz), 0.0, mz};
    }
    LLGParams params{GAMMA_E_RAD_PER_S_T, 0.0, 1e-13, {0.0, 0.0, 0.0}};
    Material plus;
    plus.ms = 8.0e5; plus.a_ex = 13e-12; plus.dmi = 1e-4;
    Material minus = plus;
    minus.dmi = -1e-4;
    VectorField2D bp(grid), bm(grid);
    effective_field::build_h_eff(grid, m, bp, params, plus);
    effective_field::build_h_eff(grid, m, bm, params, minus);
    double bxp = bp.data[4][0], bxm = bm.data[4][0];
    CHECK(std::fabs(bxp) > 0.0, "bx=%g", bxp);
    CHECK(std::fabs(bxp + bxm) < 1e-10 * std::fmax(std::fabs(bxp), 1.0), "bx+=%g bx-=%g", bxp, bxm);
}

// tests/validation.rs:646-745
static void macrospin_precession_quarter_turn_rk45_adaptive() {
    Grid2D grid(1, 1, 1.0, 1.0, 1.0);
    VectorField2D m(grid);
    m.set_uniform(1.0, 0.0, 0.0);
    double b0 = 0.1;
    LLGParams params{GAMMA_E_RAD_PER_S_T, 0.0, 1e-14, {0.0, 0.0, b0}};
    Material material;
    material.ms = 8.0e5;
    llg::RK45Scratch scratch(grid);
    double t_target = (PI / 2.0) / (params.gamma * b0), t = 0.0;
    size_t attempts = 0;
    while (t < t_target) {
        ++attempts;
        if (attempts >= 200000) { CHECK(false, "too many attempts"); return; }
        double remaining = t_target - t;
        if (params.dt > remaining) params.dt = remaining;
        auto [eps, accepted, dt_used] = llg::step_llg_rk45_recompute_field_adaptive(
            m, params, material, scratch, 1e-5, 0.8, 1e-18, 1e-11);
        (void)eps;
        if (accepted) t += dt_used;
    }
    auto v = m.data[0];
    CHECK(std::fabs(v[2]) < 0.1, "mz=%g", v[2]);
    CHECK(std::fabs(v[1]) > 0.9, "my=%g", v[1]);
    CHECK(std::fabs(v[0]) < 0.3, "mx=%g", v[0]);
}

// tests/test_demag_methods.rs:20-51 (dispatcher wiring): the dispatcher's FFT
// route equals the field builder's demag term.
static void dispatcher_fft_matches_builder() {
    Grid2D grid(16, 8, 4e-9, 4e-9, 3e-9);
    VectorField2D m(grid);
    for (size_t i = 0; i < m.data.size(); ++i) {
        double a = 0.37 * (double)i;
        m.data[i] = {std::cos(a) * 0.6, std::sin(a) * 0.6, 0.8};
    }
    Material mat = demag_only(8e5);
    VectorField2D b1(grid), b2(grid);
    effective_field::demag::compute_demag_field(grid, m, b1, mat);
    LLGParams p{GAMMA_E_RAD_PER_S_T, 0.0, 0.0, {0.0, 0.0, 0.0}};
    effective_field::build_h_eff(grid, m, b2, p, mat);
    double d = 0.0;
    for (size_t i = 0; i < m.data.size(); ++i)
        for (int c = 0; c < 3; ++c) d = std::fmax(d, std::fabs(b1.data[i][c] - b2.data[i][c]));
    CHECK(d < 1e-12, "max diff %g", d);
}

// tests/validation.rs:222-318: dE = -Ms V B_eff . dm at an interior cell of a
// Neel wall (init_bloch_wall -> init_neel_wall_x, vector_field.rs:103-126, 162-164).
static void energy_gradient_consistency_exchange_anisotropy() {
    Grid2D grid(16, 16, 5e-9, 5e-9, 5e-9);
    VectorField2D m(grid);
    const double x0 = 0.5 * (double)grid.nx * grid.dx, width = 5.0 * grid.dx;
    for (size_t j = 0; j < grid.ny; ++j)
        for (size_t i = 0; i < grid.nx; ++i) {
            double u = (((double)i + 0.5) * grid.dx - x0) / width;
            double mz = std::tanh(u), mx = 1.0 / std::cosh(u);
            double nn = std::sqrt(mx * mx + mz * mz);
            m.data[m.idx(i, j)] = {mx / nn, 0.0, mz / nn};
        }
    LLGParams params{GAMMA_E_RAD_PER_S_T, 0.02, 1e-13, {0.0, 0.0, 0.0}};
    Material material;
    material.ms = 8.0e5; material.a_ex = 13e-12; material.k_u = 500.0;
    VectorField2D b_eff(grid);
    effective_field::build_h_eff(grid, m, b_eff, params, material);
    const double e0 = energy::compute_total_energy(grid, m, material, params.b_ext);
    const size_t idx = grid.idx(grid.nx / 2, grid.ny / 2);
    const Vec3 m0 = m.data[idx];
    const double a[3] = {0.37, -0.24, 0.91};
    const double mdota = m0[0] * a[0] + m0[1] * a[1] + m0[2] * a[2];
    double dir[3] = {a[0] - mdota * m0[0], a[1] - mdota * m0[1], a[2] - mdota * m0[2]};
    const double dn = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (double &d : dir) d /= dn;
    const double eps = 1e-6;
    double v[3] = {m0[0] + eps * dir[0], m0[1] + eps * dir[1], m0[2] + eps * dir[2]};
    const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const Vec3 m1{v[0] / vn, v[1] / vn, v[2] / vn};
    VectorField2D m_pert = m;
    m_pert.data[idx] = m1;
    const double e1 = energy::compute_total_energy(grid, m_pert, material, params.b_ext);
    const Vec3 b = b_eff.data[idx];
    const double de_pred = -material.ms * grid.cell_volume() *
                           (b[0] * (m1[0] - m0[0]) + b[1] * (m1[1] - m0[1]) + b[2] * (m1[2] - m0[2]));
    const double bn = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double scale = std::fmax(material.ms * grid.cell_volume() * bn * eps, 1e-30);
    CHECK(std::fabs((e1 - e0) - de_pred) < 5e-2 * scale, "num=%g pred=%g", e1 - e0, de_pred);
}

// validation.rs:88-122 (macrospin relaxes toward the easy axis), through the
// relax controller (relax.rs:334-774) with the SP4-style threshold settings.
static void macrospin_relax_controller_reaches_easy_axis() {
    Grid2D grid(1, 1, 1.0, 1.0, 1.0);
    VectorField2D m(grid);
    m.set_uniform(0.6, 0.0, 0.8);
    LLGParams params{GAMMA_E_RAD_PER_S_T, 0.5, 1e-13, {0.0, 0.0, 0.0}};
    Material material;
    material.ms = 8.0e5; material.k_u = 5e5;
    llg::RK23Scratch scratch(grid);
    relax::RelaxSettings settings;
    auto rep = relax::relax_with_report(grid, m, params, material, scratch, FieldMask::ExchAnis, settings);
    CHECK(rep.stop_reason == relax::RelaxStopReason::TightenFloorReached, "stop=%d", (int)rep.stop_reason);
    CHECK(rep.final_torque && *rep.final_torque <= 1e-4, "torque");
    CHECK(settings.max_err == 1e-9, "max_err=%g", settings.max_err);
    CHECK(m.data[0][2] > 0.999999, "mz=%g", m.data[0][2]);
    CHECK(scratch.last_b_eff().has_value(), "last_b_eff");
    scratch.invalidate_last_b_eff();
    CHECK(!scratch.last_b_eff().has_value(), "invalidated");
}

// ovf.rs:264-316: OvfSeries dumps at t = 0, every, 2 every, ... as time passes.
static void ovf_series_dumps_on_schedule() {
    Grid2D grid(4, 3, 1e-9, 1e-9, 1e-9);
    VectorField2D m(grid);
    const std::string dir = "/tmp/llgmi_cpp_ovf";
    ovf::OvfSeries series(dir, 1e-12);
    std::vector<std::string> written;
    auto path = [&](uint64_t frame, double) {
        std::string p = dir + "/m" + std::to_string(frame) + ".ovf";
        written.push_back(p);
        return p;
    };
    series.dump_initial(grid, m, path, ovf::OvfMeta::magnetization());
    series.maybe_dump_due(grid, m, 2.5e-12, path, [](double t) { return ovf::OvfMeta::magnetization().with_total_sim_time(t); });
    CHECK(written.size() == 3, "frames=%zu", written.size());
    CHECK(series.frame() == 3, "frame=%llu", (unsigned long long)series.frame());
    std::FILE *f = std::fopen(written.back().c_str(), "rb");
    CHECK(f != nullptr, "missing %s", written.back().c_str());
    if (f) std::fclose(f);
}

int main() {
    const std::pair<const char *, std::function<void()>> tests[] = {
        {"demag_single_cell_cube_is_close_to_minus_mu0_over_3", demag_single_cell_cube_is_close_to_minus_mu0_over_3},
        {"demag_uniform_2x2_cube_cells_has_symmetry_and_reasonable_factor", demag_uniform_2x2_cube_cells_has_symmetry_and_reasonable_factor},
        {"demag_uniform_2x2_prints_nxx_nyy_nzz", demag_uniform_2x2_prints_nxx_nyy_nzz},
        {"demag_energy_is_nonnegative_for_uniform_state", demag_energy_is_nonnegative_for_uniform_state},
        {"dmi_field_flips_sign_with_d", dmi_field_flips_sign_with_d},
        {"macrospin_precession_quarter_turn_rk45_adaptive", macrospin_precession_quarter_turn_rk45_adaptive},
        {"dispatcher_fft_matches_builder", dispatcher_fft_matches_builder},
        {"energy_gradient_consistency_exchange_anisotropy", energy_gradient_consistency_exchange_anisotropy},
        {"macrospin_relax_controller_reaches_easy_axis", macrospin_relax_controller_reaches_easy_axis},
        {"ovf_series_dumps_on_schedule", ovf_series_dumps_on_schedule},
    };
    for (auto &t : tests) {
        std::printf("test %s\n", t.first);
        std::fflush(stdout);
        const auto t0 = std::chrono::steady_clock::now();
        int before = g_fail;
        try {
            t.second();
        } catch (const std::exception &e) {
            std::printf("  FAIL exception: %s\n", e.what());
            ++g_fail;
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("  %s (%.2f s)\n", g_fail == before ? "ok" : "FAILED", sec);
    }
    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "PASSED", g_fail);
    return g_fail ? 1 : 0;
}
