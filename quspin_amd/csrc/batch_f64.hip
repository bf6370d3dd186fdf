// Instantiations of the batched wavefront-per-row kernels for double vectors.
#include "batch_impl.h"

QSX_DEFINE_WAVE_LAUNCH(qsx_wave_launch_f64, double)
