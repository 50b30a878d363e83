// gr_internal.hpp -- host-side state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/gr_capi.h"

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
}  // namespace gr_detail
