// The stepping kernels in FP64 (namespace rp::f64): the same source as
// rp_kernels.hip with every real a double, for runs that must reproduce the
// reference's FP64 CPU arithmetic bit for bit.
#define RP_FP64 1
#include "rp_kernels.hip"
