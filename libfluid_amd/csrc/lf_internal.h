// lf_internal.h -- shared host-side internals of libfluid_hip.so.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "lf_abi.h"

namespace lf {

// Records a printf-style message as the calling thread's last error and
// returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Every device allocation of the library goes through dev_alloc / dev_free.
// dev_alloc counts the allocations the library holds and fails on request
// (lf_debug_fail_alloc: the test hook for allocation-failure paths);
// dev_free frees and nulls the pointer, so releasing twice is harmless.
int dev_alloc(void **p, size_t bytes);
void dev_free_raw(void *p);
void debug_fail_alloc(int64_t nth);  // nth < 0: never
int64_t debug_live_allocs();
template <class T>
void dev_free(T *&p) {
    if (p) dev_free_raw((void *)p);
    p = nullptr;
}

// createFluidBlock's lattice (fluidsystem.cu:66-127) as per-axis coordinate
// lists: the FP32 running sums of its three loops (x outermost, z innermost);
// the block is the first min(n, |ax| |ay| |az|) points of their product.
// Also returns the mass (particle volume x rho0) and particle size.
void block_axes(uint32_t n, const float mn[3], const float mx[3], float rest_density, std::vector<float> axes[3],
                float *mass, float *particle_size);

}  // namespace lf
