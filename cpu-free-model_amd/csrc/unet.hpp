// Internal interface between the warm-start network (unet.hip) and the
// batched CG plan (cg_batch.hip); not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "cfm.h"

namespace cfm {

struct UnetView {
    int n, batch;
    hipStream_t stream;
    float *input;  // the network's own input staging buffer [batch][2][n^3]
};

UnetView unet_view(cfm_unet *net);
// Forward pass over batch * 2 * n^3 device floats at d_in, on the network's
// stream (not synchronised); *d_pred receives the device output [batch][n^3].
cfm_status unet_forward_device(cfm_unet *net, const float *d_in, const float **d_pred);

}  // namespace cfm
