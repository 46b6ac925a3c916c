// lf_io.h -- checkpoint file format (internal to libfluid_hip.so).
#pragma once

#include <stdint.h>

#include <vector>

#include "lf_abi.h"

namespace lf {

// Writes / reads a checkpoint: parameters, particle count, step count, last
// substep length and the exact state (12 floats per particle in id order:
// position high and low words, velocity, acceleration), with a checksum.
int write_checkpoint(const char *path, const lf_params &p, uint32_t n, uint64_t steps, float dt,
                     const std::vector<float> &state);
int read_checkpoint(const char *path, lf_params *p, uint32_t *n, uint64_t *steps, float *dt,
                    std::vector<float> &state);

}  // namespace lf
