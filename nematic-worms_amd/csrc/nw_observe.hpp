// nw_observe.hpp -- device sums behind the order parameters (nw_observe.hip)
#pragma once

#include "nw_internal.hpp"

namespace nw {

// indices of the FP64 sums (the layout of NWObservables.sums)
enum : int {
    kObsTxx, kObsTxy, kObsTxz, kObsTyy, kObsTyz, kObsTzz,  // sum of t t over bead tangents
    kObsTx, kObsTy, kObsTz,                                // sum of t
    kObsBeads,                                             // beads with a tangent
    kObsBxx, kObsBxy, kObsBxz, kObsByy, kObsByz, kObsBzz,  // sum of b b over bond tangents
    kObsBonds,                                             // bonds
    kObsVdotT,                                             // sum of v . t
    kObsV2,                                                // sum of |v|^2
    kObsSums
};

// rows of partial sums launch_observe writes (one per workgroup)
uint32_t observe_rows(uint32_t W, uint32_t n);
// partial: observe_rows x kObsSums doubles; out: kObsSums doubles
hipError_t launch_observe(const float4 *pos, const float4 *vel, uint32_t W, double *partial, double *out,
                          const KP &kp, hipStream_t st);

}  // namespace nw
