// Device-resident solver state behind the C-ABI handle (include/nemo_gpu.h)
// and the per-degree stage launch, shared by the translation units of
// libnemo_gpu.so: nemo_gpu.hip (entry points) and stage_m{3,4,5}.hip (the
// stage kernels of one polynomial degree each, compiled in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/nemo_gpu.h"
#include "host_ops.hpp"

namespace nemo_host {

struct Solver {
    nemo_gpu_config cfg{};
    HostOps ops{};
    nemo::RKCoeffs rk{};
    int M = 0, M3 = 0, V = 5;
    long long nelem = 0;
    long long cap = 0;                 // elements the mesh-sized buffers hold (>= nelem)
    bool have_mesh = false, have_state = false;

    double* u[2] = {nullptr, nullptr};
    int cur = 0;
    double* r0 = nullptr;
    double* r1 = nullptr;
    double* rhs = nullptr;     // RHS / scratch of state size
    double* alpha = nullptr;   // [nelem]
    int32_t* nbr = nullptr;
    double* delta = nullptr;
    long long nmortar = 0;             // non-conforming (2:1) faces
    int32_t* mortars = nullptr;        // [nmortar][6]
    long long nbound = 0;              // boundary faces (non-periodic meshes)
    int32_t* bfaces = nullptr;         // [nbound][3]: element, face, kind
    long long nslots = 0;              // 5*nmortar + nbound face slots
    double* mortar_flux = nullptr;     // [nslots][V][M*M]: mortar slots, then boundary slots
    int* counter = nullptr;            // remap fallbacks
    int32_t* units = nullptr;          // replacement units of the last remesh
    long long units_cap = 0;
    double* an_part = nullptr;         // analysis partials
    long long an_cap = 0;
    double* an_out = nullptr;          // [NEMO_ANALYZE_NSUMS]
    double* scratch = nullptr;
    double* cfl_red = nullptr;         // [3 * CFL_GRID] second level of the CFL reduction
    unsigned int* cfl_ticket = nullptr;  // blocks done in the current CFL reduction         // this device's `scratch` (allocation-free analysis)
    double* dt_dev = nullptr;  // [8]: dt, lmax, vmax, glm c_h, glm alpha
    double* part = nullptr;    // per-block CFL partials [nblocks][4]
    int* status = nullptr;
    double* hist = nullptr;
    long long hist_cap = 0;
    hipStream_t stream = nullptr;

    // learned indicator
    float* mlp_p = nullptr;     // canonical parameters
    float* mlp_inf = nullptr;   // inference layout
    float* mlp_m = nullptr;     // Adam moments
    float* mlp_v = nullptr;
    unsigned long long* mlp_stats = nullptr;  // judged, overridden, fallback
    double* alpha_step = nullptr;  // [nelem]
    bool mlp_ready = false;
    uint32_t step_counter = 0;
    double sample_fraction = 0.05, judge_tol = 0.05;

    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;

    size_t state_doubles() const { return size_t(V) * size_t(nelem) * size_t(M3); }
};

// Per-launch options of the stage kernel.
struct LaunchOpts {
    int refresh = 1;            // learned indicator: compute the factor of the step
    float* feat_out = nullptr;  // training data: features of every element
    int alpha_only = 0;         // indicator only (Euler): the factor to alpha_out, no operator
};

// Stage launch of one polynomial degree (M = n_poly + 1 nodes per direction),
// defined in stage_launch.hpp and instantiated in stage_m<M>.hip.
template <int M>
int launch_stage_m(Solver& s, int stage, double dt, const double* dt_dev, double* rhs_out, double* alpha_out,
                   double* cfl_part, const LaunchOpts& o);

}  // namespace nemo_host
