// md_frames.h -- internal: what the frame writer (md_frames.hip) reads of a
// context (md_api.hip).
#pragma once

#include "md_device.h"

struct md_ctx;

namespace md {

// The current state of one replica as device data: row r of its frame file
// is slot slot0[g] + (r - row0[g]) of group g (row0[g] <= r < row0[g + 1]),
// particle id type[g]; the record buffer is rec[ctrl->cur], and ctrl->box is
// the replica's box, both read on the device when the writer's kernels run.
struct FrameSource {
    hipStream_t st;
    int device;
    const unsigned char* rec[2];
    int32_t stride;  // bytes per record; x, y, z, vx, vy, vz are its first six doubles
    const RepCtrl* ctrl;
    int32_t n_groups;
    int64_t row0[kMaxTypes + 1];
    int64_t slot0[kMaxTypes];
    int32_t type[kMaxTypes];
};

}  // namespace md

// Fills *out for replica `replica` of c (the buffers of the moment: call it
// again after anything that may reallocate them).
__attribute__((visibility("hidden"))) int md_frame_source(md_ctx* c, int replica, md::FrameSource* out);
