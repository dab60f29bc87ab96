// Stage kernels of n_poly = 3 (M = 4 nodes per direction), gfx950.
#include "stage_launch.hpp"

namespace nemo_host {
template int launch_stage_m<4>(Solver& s, int stage, double dt, const double* dt_dev, double* rhs_out,
                                double* alpha_out, double* cfl_part, const LaunchOpts& o);
}  // namespace nemo_host
