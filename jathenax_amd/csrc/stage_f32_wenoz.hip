// Fused stage kernels for float state with wenoz reconstruction (all solvers, nvar 3-5).
#include "jx_stage.hpp"

JX_STAGE_LAUNCHER(float, jx::RECON_WENOZ, jx_stage_f32_wenoz)
