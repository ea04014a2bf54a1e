// Batch calls whose device memory comes from a fixed region the caller owns (a Carve), allocating none.
#pragma once

#include <stdint.h>

#include "../../include/dinara.h"
#include "device.h"

namespace dinara {

// dinara_gpu_distances for unit costs (times a scale), global, no band: the bit-parallel path with every device
// buffer taken from `carve`, its launches made one by one on the device's stream (no graph, side streams or timing
// events). DINARA_SCRATCH_FULL when the batch needs more than the region holds; DINARA_INVALID_MODE for other costs.
int64_t carved_unit_distances(Carve *carve, int64_t gpu, int64_t pairs, const char *const *references,
                              const int64_t *reference_lengths, const char *const *queries,
                              const int64_t *query_lengths, const dinara_costs *costs, const dinara_mode *mode,
                              const dinara_options *options, int64_t threads, int64_t *results);

}  // namespace dinara
