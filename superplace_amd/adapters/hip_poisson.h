// The HIP Poisson backend as the reference's PoissonBackend
// (src/numeric/poisson_backend.h:18-45).  Pure C++, like
// src/numeric/cuda/cuda_poisson.h: it goes to src/numeric/hip/ in the
// reference tree and includes no HIP header.
#pragma once

#include <memory>

#include "numeric/poisson_backend.h"

namespace sp {

// nullptr when no GPU is present (the factory then keeps the CPU backend),
// as makeCudaPoissonBackend (src/numeric/cuda/cuda_poisson.h:11-16); throws
// std::runtime_error on any other failure.
std::unique_ptr<PoissonBackend> makeHipPoissonBackend(const BackendConfig& cfg);

}  // namespace sp
