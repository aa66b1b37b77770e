// Geometric-multigrid Poisson demag (second demag path).
#pragma once

#include "internal.hpp"

namespace llgmi {

struct MGSolver {
    MGSolver(int64_t, int64_t, double, double, double, hipStream_t, Profiler *) {}
    template <typename T>
    void demag_field(const T *, double, T *) {
        throw Error(3, "multigrid demag is not built");
    }
};

}  // namespace llgmi
