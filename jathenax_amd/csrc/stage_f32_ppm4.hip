// Fused stage kernels for float state with ppm4 reconstruction (all solvers, nvar 3-5).
#include "jx_stage.hpp"

JX_STAGE_LAUNCHER(float, jx::RECON_PPM4, jx_stage_f32_ppm4)
