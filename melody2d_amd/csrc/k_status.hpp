// Body deactivation: Update_status (Code/Proximity.h:550-1007) after each
// contact update, and the contact sends inactive bodies keep making.
#pragma once
#include "k_block.hpp"
#include "mel_device.hpp"

namespace mel {

constexpr int STATUS_THREADS = 256;

// One workgroup (STATUS_THREADS): the rules in order, then, for the bodies
// they deactivated, the last master records of their elements become
// ghost sends, their current elements and node blocks are marked inactive
// and their rigid nodes keep the force copies of their last step.
__global__ void k_update_status(DevModel m, BlockArrays B);

// Send_contact_forces of the inactive bodies: their ghost records appended
// to the border-node buckets (rigid masters read theirs in the rigid-body
// resultants)
__global__ void k_ghost_sends(DevModel m);

}  // namespace mel
