// Shared declarations of the learned edge prior (prior.cpp) used by the
// engine to turn a trained network into per-instance heuristic multipliers.
#pragma once

#include "../../include/faco_engine.h"

namespace faco {
// Device pointer of the weight image the prior's kernels read (current
// after every call that changed the parameters has returned).
const void* prior_device_image(const faco_prior* prior);
}  // namespace faco
