// Fused stage kernels for double state with ppmx reconstruction (all solvers, nvar 3-5).
#include "jx_stage.hpp"

JX_STAGE_LAUNCHER(double, jx::RECON_PPMX, jx_stage_f64_ppmx)
