// GLM-MHD direction pass of the fused stage kernel (gfx950).
//
// Same mapping as the Euler pass (a thread owns one grid line of one
// element), with the Powell/GLM non-conservative terms of the split form:
//   volume  sum_m S_im [ f#(u_i,u_m) + Phi(u_i) {{B_n, psi}}_im ]
//   faces   f*(u_-, u_+) + Phi(u_inner) {{B_n, psi}}_face
//   FV      per side of each subcell interface, Phi(u_cell) {{B_n, psi}}
// (equations/mhd/non_conservative_flux_mod.f90:10-45; placement as in
// kernel/mcfv/3d/kernel_flux_mod.f90:478-494).  Phi(u_i) is linear in the
// averages, so the volume sum is formed as Phi(u_i) . (sum_m S_im {{.}}_im)
// with two scalar accumulators per node instead of one 9-vector per pair.
#pragma once

#include "dg_common.hpp"

namespace nemo {

template <int SURF, int D, typename Real>
__device__ __forceinline__ void mhd_riemann(const MhdGas<Real>& g, const MhdState<Real>& L, const MhdState<Real>& R,
                                            Real* f) {
    if constexpr (SURF == SURF_HLLD)
        mhd_hlld<D>(g, L, R, f);
    else
        mhd_rusanov<D>(g, L, R, f);
}

template <int D, int M, int BE, int VOL, int SURF, typename Real, typename TU>
__device__ __forceinline__ void direction_pass_mhd(const StageArgs<M, Real>& a, const MhdGas<Real>& g, long long e0,
                                                   const TU (*s_u)[BE * M * M * M],
                                                   Real (*s_acc)[BE * M * M * M], const Real alpha) {
    using Sh = StageShape<M, BE>;
    constexpr int M2 = Sh::M2, M3 = Sh::M3, V = 9;
    constexpr int BN = Mag<D, 0>::idx;
    constexpr bool BETA = (VOL == VOL_EC);
    const int tid = threadIdx.x;
    const int el = tid / M2;
    const int line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const long long e = e0 + el;
    const bool valid = e < a.nelem;
    const long long EM3 = a.nelem * (long long)M3;
    const int base = el * M3;

    MhdState<Real> S[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int n = bslot<M>(base, nidx<D, M>(i, ta, tb));
#pragma unroll
        for (int h = 0; h < V; ++h) S[i].u[h] = Real(s_u[h][n]);
        mhd_make_state<BETA>(g, S[i]);
    }
    Real fm[V], fp[V];
    // Lanes past the last element read element 0's neighbours (their results
    // are discarded): unconditional loads, so later loads need not wait.
    const long long ec = valid ? e : 0;
    const Real jac = Real(2.0 / a.delta[ec * 3 + D]);
    {
        // A face on a non-conforming interface takes its flux, with this
        // side's non-conservative term, from the mortar pass: the face
        // computation then runs on the element's own face node and is
        // discarded.
        const long long em = a.nbr[ec * 6 + 2 * D];
        const long long ep = a.nbr[ec * 6 + 2 * D + 1];
        MhdState<Real> Nm, Np;
        {
            const double* um = a.u_in + (em < 0 ? ec * M3 + nidx<D, M>(0, ta, tb) : em * M3 + nidx<D, M>(M - 1, ta, tb));
            const double* up = a.u_in + (ep < 0 ? ec * M3 + nidx<D, M>(M - 1, ta, tb) : ep * M3 + nidx<D, M>(0, ta, tb));
#pragma unroll
            for (int h = 0; h < V; ++h) {
                Nm.u[h] = Real(um[h * EM3]);
                Np.u[h] = Real(up[h * EM3]);
            }
        }
        mhd_make_state<false>(g, Nm);
        mhd_make_state<false>(g, Np);
        mhd_riemann<SURF, D>(g, Nm, S[0], fm);
        mhd_riemann<SURF, D>(g, S[M - 1], Np, fp);
        mhd_nc_add<D>(g, S[0], Real(0.5) * (Nm.u[BN] + S[0].u[BN]), Real(0.5) * (Nm.u[GLMP] + S[0].u[GLMP]), fm);
        mhd_nc_add<D>(g, S[M - 1], Real(0.5) * (S[M - 1].u[BN] + Np.u[BN]),
                      Real(0.5) * (S[M - 1].u[GLMP] + Np.u[GLMP]), fp);
        if (em < 0) {
            const double* mf = a.mortar_flux + size_t(-em - 1) * V * M2 + ta + M * tb;
#pragma unroll
            for (int h = 0; h < V; ++h) fm[h] = Real(mf[h * M2]);
        }
        if (ep < 0) {
            const double* mf = a.mortar_flux + size_t(-ep - 1) * V * M2 + ta + M * tb;
#pragma unroll
            for (int h = 0; h < V; ++h) fp[h] = Real(mf[h * M2]);
        }
    }

    Real c[M][V];
    {
        Real sB[M], sG[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            sB[i] = sG[i] = Real(0);
#pragma unroll
            for (int h = 0; h < V; ++h) c[i][h] = Real(0);
        }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int m = i + 1; m < M; ++m) {
                Real F[V];
                if constexpr (VOL == VOL_EC)
                    mhd_ec_flux<D>(g, S[i], S[m], F);
                else
                    mhd_central_flux<D>(g, S[i], S[m], F);
                const Real Bavg = Real(0.5) * (S[i].u[BN] + S[m].u[BN]);
                const Real Gavg = Real(0.5) * (S[i].u[GLMP] + S[m].u[GLMP]);
                const Real sim = a.op.S[i][m], smi = a.op.S[m][i];
#pragma unroll
                for (int h = 0; h < V; ++h) {
                    c[i][h] += sim * F[h];
                    c[m][h] += smi * F[h];
                }
                sB[i] += sim * Bavg;
                sG[i] += sim * Gavg;
                sB[m] += smi * Bavg;
                sG[m] += smi * Gavg;
            }
#pragma unroll
        for (int i = 0; i < M; ++i) mhd_nc_add<D>(g, S[i], sB[i], sG[i], c[i]);
    }
#pragma unroll
    for (int h = 0; h < V; ++h) {
        c[0][h] = c[0][h] - fm[h] * a.op.inv_w[0];
        c[M - 1][h] = c[M - 1][h] + fp[h] * a.op.inv_w[M - 1];
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
#pragma unroll
        for (int h = 0; h < V; ++h) c[i][h] = -jac * c[i][h];

    // FV subcells on the LGL sub-grid with the non-conservative terms per
    // interface side: cell i gets Phi(u_i) ({{.}}_{i+1/2} - {{.}}_{i-1/2}),
    // the element faces already carry theirs in fm/fp.
    if (alpha > Real(0)) {
        const Real beta = Real(1) - alpha;
        Real Fprev[V];
#pragma unroll
        for (int h = 0; h < V; ++h) Fprev[h] = fm[h];
        Real Bl = Real(0), Gl = Real(0);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            Real Fnext[V];
            Real Br = Real(0), Gr = Real(0);
            if (i == M - 1) {
#pragma unroll
                for (int h = 0; h < V; ++h) Fnext[h] = fp[h];
            } else {
                mhd_riemann<SURF, D>(g, S[i], S[i + 1], Fnext);
                Br = Real(0.5) * (S[i].u[BN] + S[i + 1].u[BN]);
                Gr = Real(0.5) * (S[i].u[GLMP] + S[i + 1].u[GLMP]);
            }
            Real dF[V];
#pragma unroll
            for (int h = 0; h < V; ++h) dF[h] = Fnext[h] - Fprev[h];
            mhd_nc_add<D>(g, S[i], Br - Bl, Gr - Gl, dF);
            const Real sc = -jac * a.op.inv_w[i];
#pragma unroll
            for (int h = 0; h < V; ++h) {
                c[i][h] = beta * c[i][h] + alpha * (sc * dF[h]);
                Fprev[h] = Fnext[h];
            }
            Bl = Br;
            Gl = Gr;
        }
    }

#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int n = bslot<M>(base, nidx<D, M>(i, ta, tb));
#pragma unroll
        for (int h = 0; h < V; ++h) {
            if constexpr (D == 0)
                s_acc[h][n] = c[i][h];
            else
                s_acc[h][n] += c[i][h];
        }
    }
}

}  // namespace nemo
