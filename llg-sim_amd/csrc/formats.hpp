// Host-side writers and readers of llg-sim's on-disk formats (formats.hip).
#pragma once

#include "../../include/llgmi.h"

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

}  // namespace llgmi
