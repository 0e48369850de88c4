// batch_internal.h -- functions of batch.cpp used by autotune.cpp.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/atx.h"

namespace atx {

struct Batch;
Batch* batch_cast(atx_batch* h);
int batch_accel(const Batch* b);
// True when no scene has a tree with the batch's acceleration setting.
bool batch_flat_only(const Batch* b);
uint64_t batch_scene_pixels(const Batch* b, uint32_t id);
// Renders each scene's first max_frames frames; *ms_out gets the HIP-event time.
int batch_run(Batch* b, uint32_t max_frames, double* ms_out);
bool batch_is_kahan(const Batch* b);
// Restricts rendering to the given scene ids (empty = all scenes).
void batch_set_subset(Batch* b, const std::vector<uint32_t>& ids);
size_t batch_size(const Batch* b);
// Workload summary used as the autotune cache key.
void batch_signature(const Batch* b, uint64_t* scenes, uint64_t* pixels, uint64_t* spheres, uint64_t* spp,
                     uint64_t* bounces);

}  // namespace atx
