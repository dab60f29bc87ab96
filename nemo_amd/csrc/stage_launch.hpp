// Stage-kernel launches of one polynomial degree: kernel arguments from the
// solver state and the dispatch over equations, volume and surface fluxes,
// precision and the learned indicator.  Included by stage_m{3,4,5}.hip only,
// each of which instantiates launch_stage_m for its degree.
#pragma once

#include <cmath>

#include "solver.hpp"
#include "dg_mortar.hpp"

namespace nemo_host {

template <int M, typename Real>
nemo::StageArgs<M, Real> make_args(Solver& s, int stage, double dt, const double* dt_dev, double* rhs_out,
                                   double* alpha_out, double* cfl_part, const LaunchOpts& o) {
    nemo::StageArgs<M, Real> a{};
    a.mlp = s.mlp_inf;
    a.alpha_step = s.alpha_step;
    a.feat_out = o.feat_out;
    a.mlp_stats = s.mlp_stats;
    a.alpha_refresh = o.refresh;
    a.alpha_only = o.alpha_only;
    a.step_index = s.step_counter;
    a.sample_thresh = s.sample_fraction >= 1.0 ? 0xFFFFFFFFu : uint32_t(s.sample_fraction * 4294967296.0);
    a.judge_tol = s.judge_tol;
    a.u_in = s.u[s.cur];
    a.u_out = s.u[s.cur ^ 1];
    a.r0 = s.r0;
    a.r1 = s.r1;
    a.rhs_out = rhs_out;
    a.alpha_out = alpha_out;
    a.nbr = s.nbr;
    a.mortar_flux = s.mortar_flux;
    a.delta = s.delta;
    a.dt_dev = dt_dev;
    a.glm = s.dt_dev + 3;
    a.status = s.status;
    a.cfl_part = cfl_part;
    a.nelem = s.nelem;
    a.dt = dt;
    a.stage = stage;
    a.timedisc = s.cfg.timedisc;
    a.indicator = s.cfg.indicator;
    a.repair = s.cfg.repair;
    a.alpha_max = s.cfg.alpha_max;
    a.alpha_min = s.cfg.alpha_min;
    a.ind_threshold = 0.5 * std::pow(10.0, -1.8 * std::pow(double(M), 0.25));
    a.ind_s = std::log((1.0 - 0.0001) / 0.0001);
    a.gas = make_gas<Real>(s.cfg.gamma);
    a.mhd = make_mhd<Real>(s.cfg.gamma, s.cfg.mu0);
    a.ph64 = make_phys64(s.cfg);
    a.op = make_table<M, Real>(s.ops);
    a.rk = s.rk;
    for (int i = 0; i < M; ++i) a.rk.w64[i] = s.ops.w[i];
    return a;
}

template <int EQ, int M, int VOL, int SURF, typename Real>
void launch_stage_t(Solver& s, const nemo::StageArgs<M, Real>& a) {
    constexpr int BE = Tile<M>::BE;
    using Sh = nemo::StageShape<M, BE>;
    const long long nblocks = (s.nelem + BE - 1) / BE;
    hipLaunchKernelGGL((nemo::stage_kernel<EQ, M, BE, VOL, SURF, Real>), dim3(unsigned(nblocks)), dim3(Sh::NT), 0,
                       s.stream, a);
}

// Fluxes of the non-conforming faces of the stage input (before the stage kernel).
template <int M, typename Real>
void launch_mortar(Solver& s) {
    nemo::MortarArgs<M, Real> a{};
    a.u_in = s.u[s.cur];
    a.flux = s.mortar_flux;
    a.mortars = s.mortars;
    a.glm = s.dt_dev + 3;
    a.nelem = s.nelem;
    a.gas = make_gas<Real>(s.cfg.gamma);
    a.mhd = make_mhd<Real>(s.cfg.gamma, s.cfg.mu0);
    a.ph64 = make_phys64(s.cfg);
    mortar_ops(s.ops, &a.op.P[0][0][0], &a.op.R[0][0][0]);
    const dim3 grid(unsigned(s.nmortar)), block(nemo::MortarShape<M>::NT);
    if (s.cfg.equations == NEMO_EQ_GLM_MHD) {
        if (s.cfg.surface == NEMO_SURF_HLLD)
            hipLaunchKernelGGL((nemo::mortar_kernel<nemo::EQ_MHD, M, nemo::SURF_HLLD, Real>), grid, block, 0,
                               s.stream, a);
        else
            hipLaunchKernelGGL((nemo::mortar_kernel<nemo::EQ_MHD, M, nemo::SURF_RUSANOV, Real>), grid, block, 0,
                               s.stream, a);
    } else if (s.cfg.surface == NEMO_SURF_HLLC) {
        hipLaunchKernelGGL((nemo::mortar_kernel<nemo::EQ_EULER, M, nemo::SURF_HLLC, Real>), grid, block, 0, s.stream, a);
    } else {
        hipLaunchKernelGGL((nemo::mortar_kernel<nemo::EQ_EULER, M, nemo::SURF_RUSANOV, Real>), grid, block, 0, s.stream, a);
    }
}

// Fluxes of the boundary faces of the stage input (FP64 for either operator).
template <int M>
void launch_boundary(Solver& s) {
    nemo::BoundaryArgs<M> a{};
    a.u_in = s.u[s.cur];
    a.flux = s.mortar_flux;
    a.bfaces = s.bfaces;
    a.glm = s.dt_dev + 3;
    a.nelem = s.nelem;
    a.nbound = s.nbound;
    a.slot0 = 5 * s.nmortar;
    a.ph64 = make_phys64(s.cfg);
    const dim3 grid{unsigned((s.nbound * M * M + 63) / 64)}, block{64};
    if (s.cfg.equations == NEMO_EQ_GLM_MHD) {
        if (s.cfg.surface == NEMO_SURF_HLLD)
            hipLaunchKernelGGL((nemo::boundary_kernel<nemo::EQ_MHD, M, nemo::SURF_HLLD>), grid, block, 0, s.stream, a);
        else
            hipLaunchKernelGGL((nemo::boundary_kernel<nemo::EQ_MHD, M, nemo::SURF_RUSANOV>), grid, block, 0, s.stream,
                               a);
    } else if (s.cfg.surface == NEMO_SURF_HLLC) {
        hipLaunchKernelGGL((nemo::boundary_kernel<nemo::EQ_EULER, M, nemo::SURF_HLLC>), grid, block, 0, s.stream, a);
    } else {
        hipLaunchKernelGGL((nemo::boundary_kernel<nemo::EQ_EULER, M, nemo::SURF_RUSANOV>), grid, block, 0, s.stream,
                           a);
    }
}

template <int M, typename Real>
int launch_stage_mr(Solver& s, int stage, double dt, const double* dt_dev, double* rhs_out, double* alpha_out,
                    double* cfl_part, const LaunchOpts& o) {
    const auto a = make_args<M, Real>(s, stage, dt, dt_dev, rhs_out, alpha_out, cfl_part, o);
    const int vol = s.cfg.volume, surf = s.cfg.surface;
    if (s.nmortar > 0 && !o.alpha_only) launch_mortar<M, Real>(s);
    if (s.nbound > 0 && !o.alpha_only) launch_boundary<M>(s);
    if constexpr (M == 4) {
        // learned indicator, and the training-data pass of any Euler scheme
        if (s.cfg.equations == NEMO_EQ_EULER && (o.feat_out || s.cfg.indicator == NEMO_IND_MLP)) {
            constexpr int BE = Tile<M>::BE;
            const long long nblocks = (s.nelem + BE - 1) / BE;
            // the network runs where the factor is computed (first stage of
            // a step, diagnostics, training data); stages 2..5 read it back
            const bool net = o.feat_out || o.refresh || rhs_out;
            const dim3 grid{unsigned(nblocks)}, block{unsigned(BE * M * M)};
            if (o.feat_out || (vol == nemo::VOL_EC && surf == nemo::SURF_HLLC)) {
                if (net)
                    hipLaunchKernelGGL((nemo::stage_kernel<nemo::EQ_EULER, M, BE, nemo::VOL_EC, nemo::SURF_HLLC, Real,
                                                           nemo::IK_MLP>), grid, block, 0, s.stream, a);
                else
                    hipLaunchKernelGGL((nemo::stage_kernel<nemo::EQ_EULER, M, BE, nemo::VOL_EC, nemo::SURF_HLLC, Real,
                                                           nemo::IK_STORED>), grid, block, 0, s.stream, a);
            } else if (vol == nemo::VOL_EC && surf == nemo::SURF_RUSANOV) {
                if (net)
                    hipLaunchKernelGGL((nemo::stage_kernel<nemo::EQ_EULER, M, BE, nemo::VOL_EC, nemo::SURF_RUSANOV,
                                                           Real, nemo::IK_MLP>), grid, block, 0, s.stream, a);
                else
                    hipLaunchKernelGGL((nemo::stage_kernel<nemo::EQ_EULER, M, BE, nemo::VOL_EC, nemo::SURF_RUSANOV,
                                                           Real, nemo::IK_STORED>), grid, block, 0, s.stream, a);
            } else {
                return fail(NEMO_ERR_ARG, "the learned indicator is built for the EC volume flux with HLLC or Rusanov");
            }
            return NEMO_OK;
        }
    }
#define NEMO_CASE(E_, V_, S_)                          \
    if (s.cfg.equations == E_ && vol == V_ && surf == S_) { \
        launch_stage_t<E_, M, V_, S_, Real>(s, a);     \
        return NEMO_OK;                                \
    }
    NEMO_CASE(nemo::EQ_EULER, nemo::VOL_WEAK, nemo::SURF_RUSANOV)
    NEMO_CASE(nemo::EQ_EULER, nemo::VOL_WEAK, nemo::SURF_HLLC)
    NEMO_CASE(nemo::EQ_EULER, nemo::VOL_CENTRAL, nemo::SURF_RUSANOV)
    NEMO_CASE(nemo::EQ_EULER, nemo::VOL_CENTRAL, nemo::SURF_HLLC)
    NEMO_CASE(nemo::EQ_EULER, nemo::VOL_EC, nemo::SURF_RUSANOV)
    NEMO_CASE(nemo::EQ_EULER, nemo::VOL_EC, nemo::SURF_HLLC)
    NEMO_CASE(nemo::EQ_MHD, nemo::VOL_CENTRAL, nemo::SURF_RUSANOV)
    NEMO_CASE(nemo::EQ_MHD, nemo::VOL_CENTRAL, nemo::SURF_HLLD)
    NEMO_CASE(nemo::EQ_MHD, nemo::VOL_EC, nemo::SURF_RUSANOV)
    NEMO_CASE(nemo::EQ_MHD, nemo::VOL_EC, nemo::SURF_HLLD)
#undef NEMO_CASE
    return fail(NEMO_ERR_ARG, "unsupported equations/volume/surface/n_poly combination");
}

template <int M>
int launch_stage_m(Solver& s, int stage, double dt, const double* dt_dev, double* rhs_out, double* alpha_out,
                   double* cfl_part, const LaunchOpts& o) {
    if (s.cfg.precision == 1)
        return launch_stage_mr<M, float>(s, stage, dt, dt_dev, rhs_out, alpha_out, cfl_part, o);
    return launch_stage_mr<M, double>(s, stage, dt, dt_dev, rhs_out, alpha_out, cfl_part, o);
}

}  // namespace nemo_host
