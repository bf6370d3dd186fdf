// Instantiations of the batched wavefront-per-row kernels for float2 vectors.
#include "batch_impl.h"

QSX_DEFINE_WAVE_LAUNCH(qsx_wave_launch_c64, float2)
