// mpms_internal.hpp -- what the tool contact (mpmc_solver.hip) needs of an
// F-tracking solver (mpms_solver.hip) inside the library: its particle
// views, stream, lattice origin and tool-node buffers.  Not part of the C
// ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "mpms.h"
#include "mpms_kernels.hpp"

namespace mpms_internal {
bool is_fp64(const mpms_solver *s);
int32_t count(const mpms_solver *s);
int device(const mpms_solver *s);
hipStream_t stream(const mpms_solver *s);
void domain_min(const mpms_solver *s, double o[3]);
mpms::Particles<float> particles_f(mpms_solver *s);
mpms::Particles<double> particles_d(mpms_solver *s);
// The solver's tool-node buffers (xyz and velocity, 3 doubles per node),
// grown to n nodes, with the tool-node count set to n.  Synchronises the
// solver's stream when they grow.
int tool_node_buffers(mpms_solver *s, int32_t n, double **xyz, double **vel);
}  // namespace mpms_internal
