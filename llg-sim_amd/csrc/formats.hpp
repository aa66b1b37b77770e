// Host-side writers and readers of llg-sim's on-disk formats (formats.hip).
#pragma once

#include "../../include/llgmi.h"

#include <cstdint>
#include <string>
#include <vector>

namespace llgmi {

// Rust `{:.{prec}e}` formatting of an f64.
std::string rust_exp(double v, int prec);
void ensure_parent_dir(const std::string &path);

// OVF 2.0 rectangular mesh writers (ovf.rs:54-262); m is AoS f64, nx*ny cells;
// meta NULL = OvfMeta::magnetization().
void write_ovf2_text(const std::string &path, const llgmi_grid &g, const double *m, const llgmi_ovf_meta *meta);
void write_ovf2_binary4(const std::string &path, const llgmi_grid &g, const double *m, const llgmi_ovf_meta *meta);
std::string ovf_total_sim_time_line(double t_s);

// Demag kernel cache LLGDMAG v2 (demag_fft_uniform.rs:206-429).  base is the
// real-space base quadrant [(sy*nx + sx)*4 + {xx, xy, yy, zz}], sx, sy >= 0
// (periodic axes: the folded bins).
void kernel_cache_sizes(const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y, int64_t &px, int64_t &py);
std::string kernel_cache_path(const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y);
void write_kernel_cache(const std::string &path, const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y,
                        const std::vector<double> &base);
bool read_kernel_cache(const std::string &path, const llgmi_grid &g, int64_t pbc_x, int64_t pbc_y,
                       std::vector<double> &base);

// Hybrid MG correction-stencil cache LLGDKH8 (demag_poisson_mg.rs:2811-3052):
// the DeltaKernelKey fields (:2289-2374) as u64 in the file's order, then the
// four stencils [xx | xy | yy | zz], (2r+1)^2 each, then the three PPPM-phi
// stencils, which this library does not build (written as zeros: the
// reference reads such a file as one without phi correction, :572-577).
struct DkCacheKey {
    uint64_t nx, ny, dx_bits, dy_bits, dz_bits, pad_xy, n_vac_z;
    uint64_t stencil, prolong, coarse_op, iso27_alpha_bits;
    uint64_t bc, tree_theta_bits, tree_leaf, tree_max_depth;
    uint64_t pre_smooth, post_smooth, smoother, omega_bits, sor_omega_bits;
    uint64_t sigma_bits, radius_xy, delta_v_cycles;
};
// DeltaKernelKey::cache_path (:2335-2374)
std::string dk_cache_path(const DkCacheKey &k);
void write_dk_cache(const std::string &path, const DkCacheKey &k, const std::vector<double> &dk4);
// false when the file is missing, short or written for another key
bool read_dk_cache(const std::string &path, const DkCacheKey &k, std::vector<double> &dk4);

}  // namespace llgmi
