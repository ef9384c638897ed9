// The adaptive schemes of Solver_step (Code/Solver.h:98-486): Adaptive_Euler,
// Mass_Scaling and Adaptive_Mass_Scaling.  Every attempt is two half steps of
// the launch sequence; the kernels here replace the end of each half
// (Newton, the temporary predictor, Euler with dt/2, the error estimate, the
// mass-scaling factors) and Store / Restore.
#pragma once
#include "k_block.hpp"
#include "mel_device.hpp"

namespace mel {

// per-body scalars of the adaptive schemes (AD_BODY_FIELDS each)
enum AdBodyField {
    AB_TOTAL = 0,      // Body::total_error
    AB_MAX,            // Body::max_error
    AB_NODE,           // Body::node_for_max_error (-1 for rigid bodies)
    AB_MASS_MS,        // Body::mass_mass_scaling
    AB_FACTOR,         // Body::factor_mass_scaling (rigid bodies)
    AB_DFACTOR,        // Body::delta_factor_mass_scaling (rigid bodies)
    AB_MAXFACTOR,      // Body::max_factor_mass_scaling
    AB_DISP_T = 8,     // rigid x, y, r displacement_temporary
    AB_VEL_T = 11,     // rigid x, y, r velocity_temporary
    AD_BODY_FIELDS = 16
};

// results of an attempt read by the host (AD_OUT_FIELDS doubles)
enum AdOutField { AO_MAX_ERROR = 0, AO_MEAN_ERROR, AO_TOTAL_MASS, AO_MAX_MASS, AO_NAN, AD_OUT_FIELDS = 8 };

struct AdArrays {
    // per node
    double *uptx, *upty;       // temporary displacement parameters (Node::*_displacement_parameter_temporary)
    double *xerr, *yerr;       // Node::x_error, y_error
    double *enorm;             // Node::error_norm before its 1e-16 floor
    double *ide;               // Node::inverse_distance_estimator
    double *fsx, *fsy;         // Node::x_factor_mass_scaling, y_factor_mass_scaling
    double *fsn, *msm;         // Node::factor_mass_scaling, mass_mass_scaling
    // per body
    double* body;              // AD_BODY_FIELDS each
    double* body_nd;           // Body::nodal_distance
    int* body_drv;             // deformable bodies: bit 0 x driven, bit 1 y driven (drivendof[0], [4])
    double* out;               // AD_OUT_FIELDS
    int* kin_nodes;            // every node of the deformable bodies
    int n_kin;
    int mass_scaling;          // the scheme runs Compute_mass_scaling
    double target_error, max_ms, ctrl_ms, err_factor_ms, decrease_ms;
};

// one contiguous copy of Store / Restore: `elem` bytes times n (or times the
// device's contact-element count when per_contact is set)
struct CopyDesc {
    unsigned long long src, dst;
    unsigned long long n;
    unsigned elem, per_contact;
};

__global__ void k_ad_bc(DevModel m, double t);
__global__ void k_ad_gather(DevModel m, AdArrays A, const int* split_bn, int n_split, int half, double dt,
                            double t_end);
__global__ void k_ad_finish(DevModel m, BlockArrays B, AdArrays A, int half, double dt, double t_end);
__global__ void k_ad_error(DevModel m, AdArrays A);
__global__ void k_ad_body_reduce(DevModel m, AdArrays A);
__global__ void k_ad_global(DevModel m, AdArrays A);
__global__ void k_ad_copy(const CopyDesc* d, int n_desc, const int* c_count);
__global__ void k_ad_rigid_copy(const double* src, double* dst, int n_bodies);
__global__ void k_set_time(double* time, double t);

}  // namespace mel
