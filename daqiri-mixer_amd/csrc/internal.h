// internal.h -- shared helpers of the host side of libmixer (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "kernels.h"
#include "mixer.h"

namespace mixer {

int hip_status(hipError_t e);
int plan_slots(const mixer_array_cfg& c);
// the correlator layout contexts use for this geometry and tuning (XcorrPlan::Kind)
int plan_kind(const mixer_array_cfg& c);
std::vector<std::pair<int, int>> quad_items_for(int nb);
// contexts correlate straight from the heaps (no corner-turn pass) under the tuning in effect
bool plan_fused(const mixer_array_cfg& c);
std::vector<std::pair<int, int>> sblocks_for(int nb);
std::vector<int4> segs_for(int nb);
int get_plan(const mixer_array_cfg& c, XcorrPlan* p);
int get_twiddles(int ng, const float2** out);
int current_device();
mixer_tuning tuning_for(int n_elem);
bool has_tuning(int n_elem);

#define MIXER_TRY(x)                  \
    do {                              \
        int st_ = (x);                \
        if (st_ != MIXER_OK) return st_; \
    } while (0)
#define MIXER_HIP(x) MIXER_TRY(::mixer::hip_status(x))

}  // namespace mixer
