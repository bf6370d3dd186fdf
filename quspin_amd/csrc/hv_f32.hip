// Instantiations of the fused H.v kernel for float vectors.
#include "hv_impl.h"

QSX_DEFINE_HV_LAUNCH(qsx_hv_launch_f32, float)
