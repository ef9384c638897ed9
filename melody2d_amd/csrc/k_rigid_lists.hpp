// Per-rigid-body contact element lists (k_rigid_lists.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "mel_device.hpp"

namespace mel {

// elements one rigid body can be the master of (status bit 32 beyond)
constexpr int RIGID_LIST_CAP = 4096;

// after a contact update (or a restored state): the lists of DevModel
// rm_cnt / rm_list / rs_range from the current elements, queued on stream s
// collect: gather the elements per rigid master here (a proximity update
// has its chunk-offset pass clear the counters and its velocity-chain pass
// append the elements); then each list is sorted and the slave ranges found
hipError_t build_rigid_lists(const DevModel& d, hipStream_t s, bool collect);

}  // namespace mel
