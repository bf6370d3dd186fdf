// Instantiations of the fused H.v kernel for float2 vectors.
#include "hv_impl.h"

QSX_DEFINE_HV_LAUNCH(qsx_hv_launch_c64, float2)
