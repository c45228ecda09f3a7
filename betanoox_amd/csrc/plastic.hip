// plastic.hip -- J2 plasticity state on the device (filled in with the
// return-map kernels; until then the tangent assembly reports BHIP_ERR_STATE).
#include "bhip_internal.hpp"

const double *bhip_plastic_tangent(bhip_ctx *ctx) { (void)ctx; return nullptr; }
void bhip_plastic_release(bhip_ctx *ctx) { (void)ctx; }
