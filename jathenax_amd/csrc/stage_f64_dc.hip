// Fused stage kernels for double state with dc reconstruction (all solvers, nvar 3-5).
#include "jx_stage.hpp"

JX_STAGE_LAUNCHER(double, jx::RECON_DC, jx_stage_f64_dc)
