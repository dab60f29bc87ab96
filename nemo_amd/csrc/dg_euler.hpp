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
    // Lanes past the last element take element 0's entries (their results
    // are discarded): unconditional loads, so later loads need not wait.
    ElemGeom g;
    const long long ec = e < a.nelem ? e : 0;
#pragma unroll
    for (int f = 0; f < 6; ++f) g.nbr[f] = a.nbr[ec * 6 + f];
#pragma unroll
    for (int d = 0; d < 3; ++d) g.dx[d] = a.delta[ec * 3 + d];
    return g;
}

// Whether neighbour nb (a mortar slot when negative) is an element of the
// block starting at e0.
__device__ __forceinline__ bool in_block(long long nb, long long e0, long long nelem, int be) {
    return nb >= e0 && nb < e0 + be && nb < nelem;
}

// Neighbour nodes across the two faces of the thread's line in each
// direction, FP64 as stored.  They are requested right after the block's
// state, so that their latency hides behind the LDS staging and the
// indicator instead of stalling each direction pass.
struct FaceNodes {
    double m[3][5];  // across face 2D (x-, y-, z-)
    double p[3][5];  // across face 2D+1 (x+, y+, z+)
};

template <int D, int M, int BE, typename Real>
__device__ __forceinline__ void load_face_nodes(const StageArgs<M, Real>& a, long long e0, const ElemGeom& geo,
                                                FaceNodes& fn) {
    constexpr int M2 = M * M, M3 = M2 * M;
    const int tid = threadIdx.x;
    const int el = tid / M2;
    const int line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const long long e = e0 + el;
    const long long EM3 = a.nelem * (long long)M3;
    // A face on a non-conforming interface takes its flux from the mortar
    // pass: the Riemann solve then runs on the element's own face node and is
    // discarded.  Lanes past the last element read element 0 (their results
    // are discarded too); no branch, so the nodes stay in registers.
    const bool valid = e < a.nelem;
    const long long em = geo.nbr[2 * D];
    const long long ep = geo.nbr[2 * D + 1];
    const long long km = valid ? (em < 0 ? e : em) : 0;
    const long long kp = valid ? (ep < 0 ? e : ep) : 0;
    long long om = km * M3 + nidx<D, M>(em < 0 ? 0 : M - 1, ta, tb);
    long long op = kp * M3 + nidx<D, M>(ep < 0 ? M - 1 : 0, ta, tb);
    if constexpr (sizeof(Real) == 4) {
        // FP32 operator: a neighbour inside the block is read from LDS by the
        // surface pass, and its lanes all load one fixed address (a single
        // cache line for all of them).  (In FP64 the extra LDS reads and
        // selects cost more than the saved requests.)
        if (in_block(em, e0, a.nelem, BE)) om = 0;
        if (in_block(ep, e0, a.nelem, BE)) op = 0;
    }
    const double* um = a.u_in + om;
    const double* up = a.u_in + op;
#pragma unroll
    for (int h = 0; h < 5; ++h) {
        fn.m[D][h] = um[h * EM3];
        fn.p[D][h] = up[h * EM3];
    }
}

// Surface term of direction D: the two face Riemann problems of the thread's
// line and their lifting to the end nodes, +jac f*_- / w_0 and
// -jac f*_+ / w_{M-1}.  With LGL subcells the DG and FV surface terms are the
// same, so the term does not depend on the blending factor and is added to
// the accumulators before the volume passes (pass D = 0 initialises them).
template <int D, int M, int BE, int SURF, typename Real, typename TU>
__device__ __forceinline__ void surface_pass(const StageArgs<M, Real>& a, long long e0, const TU (*s_u)[BE * M * M * M],
                                             Real (*s_acc)[BE * M * M * M], const double* s_p64, const ElemGeom& geo,
                                             const FaceNodes& fn) {
    using Sh = StageShape<M, BE>;
    constexpr int M2 = Sh::M2, M3 = Sh::M3;
    constexpr bool F32 = sizeof(Real) == 4;
    const int tid = threadIdx.x;
    const int el = tid / M2;
    const int line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const long long e = e0 + el;
    const bool valid = e < a.nelem;
    const Gas<Real>& g = a.gas;
    const int base = el * M3;
    const int n0 = bslot<M>(base, nidx<D, M>(0, ta, tb));
    const int n1 = bslot<M>(base, nidx<D, M>(M - 1, ta, tb));

    NodeState<Real> S0, S1, Nm, Np;
#pragma unroll
    for (int h = 0; h < 5; ++h) {
        S0.u[h] = Real(s_u[h][n0]);
        S1.u[h] = Real(s_u[h][n1]);
        Nm.u[h] = Real(fn.m[D][h]);
        Np.u[h] = Real(fn.p[D][h]);
    }
    if constexpr (F32) {
        S0.p64 = s_p64[n0];
        S1.p64 = s_p64[n1];
        Nm.p64 = pressure(a.ph64.gas, fn.m[D]);
        Np.p64 = pressure(a.ph64.gas, fn.p[D]);
        // neighbours inside the block: their staged nodes in LDS
        const long long em = geo.nbr[2 * D], ep = geo.nbr[2 * D + 1];
        const bool im = valid && in_block(em, e0, a.nelem, BE);
        const bool ip = valid && in_block(ep, e0, a.nelem, BE);
        const int nm = bslot<M>((im ? int(em - e0) : 0) * M3, nidx<D, M>(M - 1, ta, tb));
        const int np = bslot<M>((ip ? int(ep - e0) : 0) * M3, nidx<D, M>(0, ta, tb));
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            const Real lm = Real(s_u[h][nm]), lp = Real(s_u[h][np]);
            Nm.u[h] = im ? lm : Nm.u[h];
            Np.u[h] = ip ? lp : Np.u[h];
        }
        const double qm = s_p64[nm], qp = s_p64[np];
        Nm.p64 = im ? qm : Nm.p64;
        Np.p64 = ip ? qp : Np.p64;
    }
    make_state<D, false>(g, S0);
    make_state<D, false>(g, S1);
    make_state<D, false>(g, Nm);
    make_state<D, false>(g, Np);
    Real fm[5], fp[5];
    riemann<SURF, D>(g, Nm, S0, fm);
    riemann<SURF, D>(g, S1, Np, fp);
    Real jac = Real(0);
    if (valid) {
        const long long em = geo.nbr[2 * D];
        const long long ep = geo.nbr[2 * D + 1];
        jac = Real(2.0 / geo.dx[D]);
        if (em < 0) {
            const double* mf = a.mortar_flux + size_t(-em - 1) * 5 * M2 + ta + M * tb;
#pragma unroll
            for (int h = 0; h < 5; ++h) fm[h] = Real(mf[h * M2]);
        }
        if (ep < 0) {
            const double* mf = a.mortar_flux + size_t(-ep - 1) * 5 * M2 + ta + M * tb;
#pragma unroll
            for (int h = 0; h < 5; ++h) fp[h] = Real(mf[h * M2]);
        }
    }
    const Real cm = jac * a.op.inv_w[0];
    const Real cp = -jac * a.op.inv_w[M - 1];
    if constexpr (D == 0) {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int n = bslot<M>(base, nidx<D, M>(i, ta, tb));
#pragma unroll
            for (int h = 0; h < 5; ++h) s_acc[h][n] = i == 0 ? cm * fm[h] : (i == M - 1 ? cp * fp[h] : Real(0));
        }
    } else {
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            s_acc[h][n0] += cm * fm[h];
            s_acc[h][n1] += cp * fp[h];
        }
    }
}

// LSRK update done by the last volume pass itself (x last): the register
// values of the thread's x-line nodes and the time step in, the new state
// of those nodes out.
template <int M>
struct XUpdate {
    double dt;
    double r0[5][M];
    double unew[5][M];
};

// Volume pass of direction D: the thread's M line nodes become node states
// once; the M(M-1)/2 two-point fluxes (or the weak-form derivative) and the
// M-1 interior FV interface fluxes read those states.  Only LDS is read.
// With XUPD and fuse set (FP64, LSRK, pass x run last) the pass does not
// store its contributions: it adds them to the accumulated surface and y/z
// terms and applies the LSRK probe to its own line's nodes (register out to
// HBM, new state into xu), with the same arithmetic as the separate update.
// Without fuse it stores them like the other passes.  NEAR1: the caller has
// found every ratio of two densities and of two betas of the block in
// [sqrt(1/2), sqrt(2)) (ec_flux).
template <int D, int M, int BE, int VOL, int SURF, typename Real, typename TU, bool XUPD = false, bool NEAR1 = false>
__device__ __forceinline__ void direction_pass(const StageArgs<M, Real>& a, long long e0, const TU (*s_u)[BE * M * M * M],
                                               Real (*s_acc)[BE * M * M * M], const Real alpha,
                                               const double* s_p64, const ElemGeom& geo,
                                               XUpdate<M>* xu = nullptr, bool fuse = false) {
    using Sh = StageShape<M, BE>;
    constexpr int M2 = Sh::M2, M3 = Sh::M3;
    constexpr bool BETA = (VOL == VOL_EC);
    const int tid = threadIdx.x;
    const int el = tid / M2;
    const int line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const long long e = e0 + el;
    const bool valid = e < a.nelem;
    const Gas<Real>& g = a.gas;
    const int base = el * M3;

    constexpr bool F32 = sizeof(Real) == 4;
    NodeState<Real> S[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int n = bslot<M>(base, nidx<D, M>(i, ta, tb));
#pragma unroll
        for (int h = 0; h < 5; ++h) S[i].u[h] = Real(s_u[h][n]);
        make_state<D, BETA>(g, S[i]);
        if constexpr (F32) S[i].p64 = s_p64[n];
    }
    Real jac = Real(0), sdx = Real(0);
    if (valid) {
        const double dx = geo.dx[D];
        jac = Real(2.0 / dx);
        sdx = Real(double(M) / dx);
    }

    Real c[M][5];
    if constexpr (VOL == VOL_WEAK) {
        // 3-D LGL extension of kernel/dgsem/weakform/2d/kernel_flux_mod.f90:242-277:
        // rhs_i = sum_j dgDiff_ij (sdx f_j) here; the surface lifting
        // lagrM_i sdx f*_- - lagrP_i sdx f*_+ = jac (d_i0 f*_- - d_iN f*_+)/w_i
        // is the surface pass's term.
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
                c[i][h] = s;
            }
    } else {
        // Split-form flux differencing (Gassner/Winters/Kopriva 2016):
        // rhs_i = -jac [ sum_m S_im f#(u_i,u_m) + (d_iN f*_+ - d_i0 f*_-)/w_i ],
        // the bracket's surface part being the surface pass's term; each pair
        // is evaluated once and used by both nodes (f# symmetric).
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
                    ec_flux<D, NEAR1>(g, S[i], S[m], F);
                else
                    central_flux<D>(S[i], S[m], F);
#pragma unroll
                for (int h = 0; h < 5; ++h) {
                    c[i][h] += a.op.S[i][m] * F[h];
                    c[m][h] += a.op.S[m][i] * F[h];
                }
            }
#pragma unroll
        for (int i = 0; i < M; ++i)
#pragma unroll
            for (int h = 0; h < 5; ++h) c[i][h] = -jac * c[i][h];
    }

    // FV subcells on the LGL sub-grid (Hennemann et al. 2021), first order:
    // rhs_i = -jac (F_{i+1/2} - F_{i-1/2}) / w_i, blended as
    // (1-alpha) DG + alpha FV; the element-face fluxes F_{-1/2} = f*_- and
    // F_{M-1/2} = f*_+ enter through the surface pass, whose term is the
    // same for DG and FV.
    if (alpha > Real(0)) {
        const Real beta = Real(1) - alpha;
        Real Fprev[5];
#pragma unroll
        for (int h = 0; h < 5; ++h) Fprev[h] = Real(0);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            Real Fnext[5];
            if (i == M - 1) {
#pragma unroll
                for (int h = 0; h < 5; ++h) Fnext[h] = Real(0);
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

    static_assert(!XUPD || sizeof(Real) == 8, "the fused update is the FP64 path's");
    bool fused = false;
    if constexpr (XUPD) fused = fuse;
    if (fused) {
        // low-storage/4-5/timedisc_probe_mod.f90:31-37, as probe_lsrk
        const int s = a.stage;
        const long long EM3 = a.nelem * (long long)M3;
        double regs[5][M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int n = bslot<M>(base, nidx<D, M>(i, ta, tb));
#pragma unroll
            for (int h = 0; h < 5; ++h) {
                const double rhs = s_acc[h][n] + c[i][h];
                const double reg = (s == 1) ? rhs : a.rk.lsrk_a[s - 1] * xu->r0[h][i] + rhs;
                regs[h][i] = reg;
                xu->unew[h][i] = S[i].u[h] + a.rk.lsrk_c[s - 1] * xu->dt * reg;
            }
        }
        // the register of the line's nodes, four consecutive doubles at N = 3
        // (32-byte aligned): two 16-byte stores per variable
        if (valid) {
            if constexpr (D == 0 && M == 4) {
                const long long g0 = e * M3 + nidx<D, M>(0, ta, tb);
#pragma unroll
                for (int h = 0; h < 5; ++h) {
                    double2* dst = reinterpret_cast<double2*>(a.r0 + h * EM3 + g0);
                    dst[0] = make_double2(regs[h][0], regs[h][1]);
                    dst[1] = make_double2(regs[h][2], regs[h][3]);
                }
            } else {
#pragma unroll
                for (int i = 0; i < M; ++i)
#pragma unroll
                    for (int h = 0; h < 5; ++h) a.r0[h * EM3 + e * M3 + nidx<D, M>(i, ta, tb)] = regs[h][i];
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int n = bslot<M>(base, nidx<D, M>(i, ta, tb));
#pragma unroll
            for (int h = 0; h < 5; ++h) s_acc[h][n] += c[i][h];
        }
    }
}

}  // namespace nemo
