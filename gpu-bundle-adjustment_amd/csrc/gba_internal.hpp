// gba_internal.hpp -- host-side internals shared by the library's translation units.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

#include "gba_c.h"

namespace gba_amd {

// Records msg as the thread's last error; returns 1 (the C ABI error code).
int set_error(const std::string& msg);

// Host camera model with the reader's semantics (zero-order branch for
// theta <= 1e-12), reference src/bal_reader.cpp:100-128.
void snavely_residual_host(const double* cam9, const double* X3, const double* obs2, double* r2);

int synth_generate(const gba_synth_spec& spec, uint64_t seed, gba_problem_view* out);
int synth_launch_bounds(int C, int P, uint64_t seed, gba_problem_view* out);

// Library exceptions carry file:line, like the reference's CUDA check macro
// (src/lm.cu:42-49); the C ABI turns them into error codes.
struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace gba_amd
