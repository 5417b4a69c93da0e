// Fused stage kernels for double state with ppm4 reconstruction (all solvers, nvar 3-5).
#include "jx_stage.hpp"

JX_STAGE_LAUNCHER(double, jx::RECON_PPM4, jx_stage_f64_ppm4)
