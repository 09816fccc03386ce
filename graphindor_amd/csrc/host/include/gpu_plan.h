// Flattening of a Schedule_IEP into the C-ABI plan description; the field
// mapping of GPUSchedule::create_from_schedule
// (gpu/component/gpu_schedule.cuh:89-190): IEP vertex rows are the first
// |in_exclusion_optimize_vertex_id| entries of the flag/coef vectors and the
// IEP positions the first |in_exclusion_optimize_coef| entries of ans_pos.
#pragma once
#include <cstdint>
#include <vector>

#include "graphindor_amd.h"
#include "schedule_IEP.h"

struct FlatPlan {
    std::vector<int32_t> adj, father, last, next, brk, loopset, rlast, rnext, rindex;
    std::vector<uint8_t> only;
    std::vector<int32_t> iep_id, iep_vcoef, iep_coef, iep_pos;
    std::vector<uint8_t> iep_vflag, iep_flag;
    gm_plan_desc desc;
};

void flatten_schedule(const Schedule_IEP& s, int mode, FlatPlan& fp);
