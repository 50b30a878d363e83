// gr_internal.hpp -- host-side state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gr_capi.h"
#include "gr_kernels.hpp"

struct gr_timer {
    hipEvent_t start = nullptr, stop = nullptr;
    bool recorded = false;
};

struct gr_context {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    gr_global_args args{};
    bool has_args = false;
    gr_tile_list_args tl{};
    bool has_tl = false;
    uint32_t* batch_counts = nullptr;  // build scratch: per-batch counts + total
    size_t batch_cap = 0;
    uint32_t* pinned = nullptr;        // host-visible copy of the total
};

namespace gr_detail {
// Records the message returned by gr_last_error() on this thread.
void set_last_error(const std::string& msg);
// The message gr_last_error() returns on this thread.
std::string get_last_error();
}  // namespace gr_detail

// Camera block of the C-ABI -> kernel camera arguments.
inline gr::CameraArgs camera_args(const gr_camera_data& c) {
    gr::CameraArgs a;
    a.view = gr::load_m4(c.view);
    a.fov_cotangent = gr::v2{c.fov_cotangent[0], c.fov_cotangent[1]};
    a.depth_scale_bias = gr::v2{c.depth_scale_bias[0], c.depth_scale_bias[1]};
    a.position = gr::v3{c.position[0], c.position[1], c.position[2]};
    a.aspect = c.aspect;
    return a;
}
