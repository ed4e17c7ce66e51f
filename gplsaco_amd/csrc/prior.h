// Shared declarations of the learned edge prior (prior.cpp) used by the
// engine to turn a trained network into per-instance heuristic multipliers.
#pragma once

#include "../../include/faco_engine.h"

namespace faco {
// Device pointer of the FP32 master parameters of a faco_prior.
const float* prior_device_params(const faco_prior* prior);
}  // namespace faco
