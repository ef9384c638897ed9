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

// Monitoring (Code/Monitoring.h:8-559): one row per step, written at row
// *n_rows of `rows` (stride doubles: the value count, then the values)
struct MonitorArgs {
    double* rows;
    int* n_rows;
    int stride, cap;
    const double* items;        // MONITORING entries, RULE_STRIDE each
    int n_items;
    int energy;                 // MONITOR_ENERGY
    const double* body_energy;  // per body: kinetic, kinetic from parameters, boundary work (k_body_energy)
    const double* stress;       // the last nodal stresses (field-major), or null before the first
    double* iter;               // the iteration number the rows carry (advanced per row)
    double neval, max_error, mean_error;
};
__global__ void k_monitor(DevModel m, BlockArrays B, MonitorArgs a);

}  // namespace mel
