// Euler direction pass of the fused stage kernel (gfx950).
#pragma once

#include "dg_common.hpp"

namespace nemo {

template <int SURF, int D, typename Real>
__device__ __forceinline__ void riemann(const Gas<Real>& g, const NodeState<Real>& uL, const NodeState<Real>& uR,
                                        Real* f) {
    if constexpr (SURF == SURF_HLLC)
        riemann_hllc<D>(g, uL, uR, f);
    else
        riemann_rusanov<D>(g, uL, uR, f);
}

// Neighbour ids and edge lengths of the thread's element, requested once per
// stage together with the state, so that the direction passes can issue
// their face loads without first waiting for the connectivity.
struct ElemGeom {
    int32_t nbr[6];
    double dx[3];
};

template <int M, typename Real>
__device__ __forceinline__ ElemGeom load_geom(const StageArgs<M, Real>& a, long long e) {
    ElemGeom g;
    const bool valid = e < a.nelem;
#pragma unroll
    for (int f = 0; f < 6; ++f) g.nbr[f] = valid ? a.nbr[e * 6 + f] : 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) g.dx[d] = valid ? a.delta[e * 3 + d] : 1.0;
    return g;
}

// One direction pass: contributions of direction D to the node accumulators.
// The thread's M line nodes (and the two neighbour face nodes) become node
// states once; the surface Riemann problems, the M(M-1)/2 two-point fluxes
// and the M-1 FV interface fluxes all read those states.
template <int D, int M, int BE, int VOL, int SURF, typename Real, typename TU>
__device__ __forceinline__ void direction_pass(const StageArgs<M, Real>& a, long long e0, const TU (*s_u)[BE * M * M * M],
                                               Real (*s_acc)[BE * M * M * M], const Real alpha,
                                               const double* s_p64, const ElemGeom& geo) {
    using Sh = StageShape<M, BE>;
    constexpr int M2 = Sh::M2, M3 = Sh::M3;
    constexpr bool BETA = (VOL == VOL_EC);
    const int tid = threadIdx.x;
    const int el = tid / M2;
    const int line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const long long e = e0 + el;
    const bool valid = e < a.nelem;
    const long long EM3 = a.nelem * (long long)M3;
    const Gas<Real>& g = a.gas;
    const int base = el * M3;

    constexpr bool F32 = sizeof(Real) == 4;
    NodeState<Real> S[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int n = base + swz<M>(nidx<D, M>(i, ta, tb));
#pragma unroll
        for (int h = 0; h < 5; ++h) S[i].u[h] = Real(s_u[h][n]);
        make_state<D, BETA>(g, S[i]);
        if constexpr (F32) S[i].p64 = s_p64[n];
    }
    Real jac = Real(0), sdx = Real(0);
    Real fm[5], fp[5];
    {
        NodeState<Real> Nm, Np;
        // faces on a non-conforming interface take their flux from the
        // mortar pass; the Riemann solve below then runs on the element's
        // own face node and is discarded
        const double* mort_m = nullptr;
        const double* mort_p = nullptr;
        if (valid) {
            const long long em = geo.nbr[2 * D];
            const long long ep = geo.nbr[2 * D + 1];
            const double dx = geo.dx[D];
            jac = Real(2.0 / dx);
            sdx = Real(double(M) / dx);
            if (em < 0) mort_m = a.mortar_flux + size_t(-em - 1) * 5 * M2 + ta + M * tb;
            if (ep < 0) mort_p = a.mortar_flux + size_t(-ep - 1) * 5 * M2 + ta + M * tb;
            const double* um = a.u_in + (em < 0 ? e : em) * M3 + nidx<D, M>(em < 0 ? 0 : M - 1, ta, tb);
            const double* up = a.u_in + (ep < 0 ? e : ep) * M3 + nidx<D, M>(ep < 0 ? M - 1 : 0, ta, tb);
            double dm[5], dp[5];
#pragma unroll
            for (int h = 0; h < 5; ++h) {
                dm[h] = um[h * EM3];
                dp[h] = up[h * EM3];
                Nm.u[h] = Real(dm[h]);
                Np.u[h] = Real(dp[h]);
            }
            if constexpr (F32) {
                Nm.p64 = pressure(a.ph64.gas, dm);
                Np.p64 = pressure(a.ph64.gas, dp);
            }
        } else {
#pragma unroll
            for (int h = 0; h < 5; ++h) {
                Nm.u[h] = S[0].u[h];
                Np.u[h] = S[M - 1].u[h];
            }
            Nm.p64 = S[0].p64;
            Np.p64 = S[M - 1].p64;
        }
        make_state<D, false>(g, Nm);
        make_state<D, false>(g, Np);
        riemann<SURF, D>(g, Nm, S[0], fm);
        riemann<SURF, D>(g, S[M - 1], Np, fp);
        if (mort_m)
#pragma unroll
            for (int h = 0; h < 5; ++h) fm[h] = Real(mort_m[h * M2]);
        if (mort_p)
#pragma unroll
            for (int h = 0; h < 5; ++h) fp[h] = Real(mort_p[h * M2]);
    }

    Real c[M][5];
    if constexpr (VOL == VOL_WEAK) {
        // 3-D LGL extension of kernel/dgsem/weakform/2d/kernel_flux_mod.f90:242-277:
        // rhs_i = lagrM_i*(sdx f*_-) - lagrP_i*(sdx f*_+) + sum_j dgDiff_ij (sdx f_j)
        Real fx[M][5];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            flux_s<D>(S[j], fx[j]);
#pragma unroll
            for (int h = 0; h < 5; ++h) fx[j][h] = sdx * fx[j][h];
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int h = 0; h < 5; ++h) {
                Real s = Real(0);
#pragma unroll
                for (int j = 0; j < M; ++j) s += a.op.dgDiff[i][j] * fx[j][h];
                Real surf = Real(0);
                if (i == 0) surf = a.op.lagrM[0] * (sdx * fm[h]);
                if (i == M - 1) surf = surf - a.op.lagrP[M - 1] * (sdx * fp[h]);
                c[i][h] = surf + s;
            }
    } else {
        // Split-form flux differencing (Gassner/Winters/Kopriva 2016):
        // rhs_i = -jac [ sum_m S_im f#(u_i,u_m) + (d_iN f*_+ - d_i0 f*_-)/w_i ];
        // each pair is evaluated once and used by both nodes (f# symmetric).
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int h = 0; h < 5; ++h) c[i][h] = Real(0);
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int m = i + 1; m < M; ++m) {
                Real F[5];
                if constexpr (VOL == VOL_EC)
                    ec_flux<D>(g, S[i], S[m], F);
                else
                    central_flux<D>(S[i], S[m], F);
#pragma unroll
                for (int h = 0; h < 5; ++h) {
                    c[i][h] += a.op.S[i][m] * F[h];
                    c[m][h] += a.op.S[m][i] * F[h];
                }
            }
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            c[0][h] = c[0][h] - fm[h] * a.op.inv_w[0];
            c[M - 1][h] = c[M - 1][h] + fp[h] * a.op.inv_w[M - 1];
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int h = 0; h < 5; ++h) c[i][h] = -jac * c[i][h];
    }

    // FV subcells on the LGL sub-grid (Hennemann et al. 2021), first order:
    // rhs_i = -jac (F_{i+1/2} - F_{i-1/2}) / w_i with the element-face
    // fluxes f*_- and f*_+ at the ends; blended as (1-alpha) DG + alpha FV.
    if (alpha > Real(0)) {
        const Real beta = Real(1) - alpha;
        Real Fprev[5];
#pragma unroll
        for (int h = 0; h < 5; ++h) Fprev[h] = fm[h];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            Real Fnext[5];
            if (i == M - 1) {
#pragma unroll
                for (int h = 0; h < 5; ++h) Fnext[h] = fp[h];
            } else {
                riemann<SURF, D>(g, S[i], S[i + 1], Fnext);
            }
            const Real sc = -jac * a.op.inv_w[i];
#pragma unroll
            for (int h = 0; h < 5; ++h) {
                c[i][h] = beta * c[i][h] + alpha * (sc * (Fnext[h] - Fprev[h]));
                Fprev[h] = Fnext[h];
            }
        }
    }

#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int n = base + swz<M>(nidx<D, M>(i, ta, tb));
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            if constexpr (D == 0)
                s_acc[h][n] = c[i][h];
            else
                s_acc[h][n] += c[i][h];
        }
    }
}

}  // namespace nemo
