// mrp_mlp_bf16.hip - BF16 MFMA MRPNN network (placeholder until the MFMA kernel lands).
#include <hip/hip_runtime.h>

#include "mrp_internal.h"
#include "mrp_mlp.h"

namespace mrp {

int pack_mlp_bf16(const float* weights_host, void** packed_dev) {
    (void)weights_host;
    *packed_dev = nullptr;
    return MRP_OK;
}

int launch_mlp_bf16(const void*, const float*, const float*, float, vec3, uint32_t, float*, hipStream_t) {
    return fail(MRP_E_STATE, "BF16 MFMA network not available");
}

}  // namespace mrp
