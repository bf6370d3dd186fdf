// Instantiations of the fused H.v kernel for double2 vectors.
#include "hv_impl.h"

QSX_DEFINE_HV_LAUNCH(qsx_hv_launch_c128, double2)
