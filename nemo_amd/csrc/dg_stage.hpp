// Fused DGSEM/FV stage kernel for gfx950.
//
// One launch per Runge-Kutta stage does what the reference spreads over the
// pre-stage face pass and the per-quad rkstep callback
// (kernel/dgsem/weakform/2d/kernel_hooks_mod.f90:15-29,
//  timedisc/rungekutta/low-storage/4-5/timedisc_hooks_mod.f90:19-56,
//  kernel/mcfv/3d/kernel_evolve_mod.f90:7-39):
//
//   surface Riemann fluxes  ->  volume integral (weak form or two-point flux
//   differencing)  ->  FV subcell fluxes where the indicator blends  ->
//   RK probe  ->  positivity repair  ->  register update.
//
// The two-phase order of the reference (all faces from the stage's input
// state, then in-place update) is kept by reading u_in and writing u_out
// (ping-pong buffers): every face is solved from the stage's input state by
// both elements that share it, so no face buffer goes through HBM.
//
// Mapping: a workgroup owns BE consecutive elements (device layout
// u[var][elem][node], node = i + M*j + M*M*k).  Their states are staged in
// LDS.  In each direction pass a thread owns one grid line of one element
// (M*M lines per element, BE*M*M threads): it solves the two face Riemann
// problems at the line's ends, evaluates the M(M-1)/2 two-point fluxes of
// the line once each (symmetric use), the M-1 interior FV fluxes, and adds
// its line's contribution to the node accumulators in LDS.  Lines of one
// direction are disjoint, so the accumulation needs no atomics; passes are
// separated by barriers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "physics.hpp"

namespace nemo {

enum { VOL_WEAK = 0, VOL_CENTRAL = 1, VOL_EC = 2 };
enum { SURF_RUSANOV = 0, SURF_HLLC = 1 };
enum { TD_LSRK54 = 0, TD_SSPRK54 = 1 };
enum { IND_NONE = 0, IND_PP = 1 };

// Operator tables for M LGL nodes (built on the host, passed by value).
template <int M, typename Real>
struct OpTable {
    Real w[M], inv_w[M];
    Real S[M][M];       // split-form matrix 2D - W^-1 B (zero diagonal)
    Real dgDiff[M][M];  // weak-form volume matrix, reference scaling
    Real lagrM[M], lagrP[M];
    Real Vinv[M][M];    // nodal -> modal (normalised Legendre)
};

struct RKCoeffs {
    double lsrk_a[5], lsrk_c[5];
    double ssp[23];  // A0..A3, B0..B3, C0..C3, D0..D3, E0..E6
    double w64[8];   // FP64 quadrature weights (repair means)
};

template <int M, typename Real>
struct StageArgs {
    const double* __restrict__ u_in;
    double* __restrict__ u_out;
    double* __restrict__ r0;
    double* __restrict__ r1;
    double* __restrict__ rhs_out;    // non-null: write the RHS, no update
    double* __restrict__ alpha_out;  // optional per-element blending factor
    const int32_t* __restrict__ nbr; // [nelem][6]
    const double* __restrict__ delta;// [nelem][3]
    const double* __restrict__ dt_dev;
    int* __restrict__ status;
    double* __restrict__ cfl_part;   // non-null: per-block CFL partials of the new state
    long long nelem;
    double dt;
    int stage;      // 1..5
    int timedisc;
    int indicator;
    int repair;
    double alpha_max, alpha_min, ind_threshold, ind_s;
    Gas<Real> gas;
    Gas<double> gas64;  // FP64 constants (CFL, FP64 pressures of the FP32 operator)
    OpTable<M, Real> op;
    RKCoeffs rk;
};

template <int D, int M>
__device__ __forceinline__ constexpr int nidx(int l, int a, int b) {
    return D == 0 ? l + M * a + M * M * b : (D == 1 ? a + M * l + M * M * b : a + M * b + M * M * l);
}

// LDS slot of node n = i + M*j + M*M*k inside an element.  For M = 4 the
// slot is 16k + 4(i^j) + (i^k): a bijection under which the 64 lanes of any
// direction pass (one node per line, 16 lines x 4 elements) read 16 distinct
// 8-byte bank pairs, 4 lanes each -- the minimum of 4 LDS passes for a
// 64-lane b64 access.  Without it the x- and y-passes hit only 4 bank pairs
// (16-way conflicts).  Other M keep the natural order.
template <int M>
__device__ __forceinline__ constexpr int swz(int n) {
    if constexpr (M == 4) {
        const int i = n & 3, j = (n >> 2) & 3, k = n >> 4;
        return 16 * k + 4 * (i ^ j) + (i ^ k);
    } else {
        return n;
    }
}

// LDS slot of block-local node index idx = element*M^3 + n.
template <int M>
__device__ __forceinline__ constexpr int slot(int idx) {
    constexpr int M3 = M * M * M;
    return (idx / M3) * M3 + swz<M>(idx % M3);
}

template <int SURF, int D, typename Real>
__device__ __forceinline__ void riemann(const Gas<Real>& g, const NodeState<Real>& uL, const NodeState<Real>& uR,
                                        Real* f) {
    if constexpr (SURF == SURF_HLLC)
        riemann_hllc<D>(g, uL, uR, f);
    else
        riemann_rusanov<D>(g, uL, uR, f);
}

// Block index -> element block, so that each XCD (blocks are dealt to the 8
// XCDs round-robin) walks a contiguous stretch of the element order and finds
// face neighbours in its own L2.
__device__ __forceinline__ long long xcd_block(long long nblocks) {
    const long long bid = blockIdx.x;
    if ((nblocks & 7) != 0) return bid;
    const long long per = nblocks >> 3;
    return (bid & 7) * per + (bid >> 3);
}

template <int M, int BE>
struct StageShape {
    static constexpr int M2 = M * M, M3 = M2 * M;
    static constexpr int NT = BE * M2;  // threads per workgroup
    static constexpr int NN = BE * M3;  // nodes per workgroup
};

// One direction pass: contributions of direction D to the node accumulators.
// The thread's M line nodes (and the two neighbour face nodes) become node
// states once; the surface Riemann problems, the M(M-1)/2 two-point fluxes
// and the M-1 FV interface fluxes all read those states.
template <int D, int M, int BE, int VOL, int SURF, typename Real>
__device__ __forceinline__ void direction_pass(const StageArgs<M, Real>& a, long long e0, double (*s_u)[BE * M * M * M],
                                               Real (*s_acc)[BE * M * M * M], const Real* s_alpha,
                                               const double* s_p64) {
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
        if (valid) {
            const long long em = a.nbr[e * 6 + 2 * D];
            const long long ep = a.nbr[e * 6 + 2 * D + 1];
            const double dx = a.delta[e * 3 + D];
            jac = Real(2.0 / dx);
            sdx = Real(double(M) / dx);
            const double* um = a.u_in + em * M3 + nidx<D, M>(M - 1, ta, tb);
            const double* up = a.u_in + ep * M3 + nidx<D, M>(0, ta, tb);
            double dm[5], dp[5];
#pragma unroll
            for (int h = 0; h < 5; ++h) {
                dm[h] = um[h * EM3];
                dp[h] = up[h * EM3];
                Nm.u[h] = Real(dm[h]);
                Np.u[h] = Real(dp[h]);
            }
            if constexpr (F32) {
                Nm.p64 = pressure(a.gas64, dm);
                Np.p64 = pressure(a.gas64, dp);
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
    const Real alpha = s_alpha[el];
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

// Persson-Peraire indicator on rho*p (Hennemann et al. 2021, sec. 3): modal
// energies of the normalised-Legendre expansion (utils/interpol_mod.f90:
// 274-311, 584-601), alpha from the logistic ramp around the threshold
// 0.5*10^(-1.8*M^0.25), clipped by alpha_min/alpha_max.
template <int M, int BE, typename Real>
__device__ __forceinline__ void indicator_pp(const StageArgs<M, Real>& a, long long e0, double (*s_u)[BE * M * M * M],
                                             Real (*s_acc)[BE * M * M * M], Real* s_alpha, const double* s_p64) {
    using Sh = StageShape<M, BE>;
    constexpr int M2 = Sh::M2, M3 = Sh::M3, NN = Sh::NN, NT = Sh::NT;
    const int tid = threadIdx.x;
    Real* q = s_acc[2];
    Real* t1 = s_acc[0];
    Real* t2 = s_acc[1];
    for (int idx = tid; idx < NN; idx += NT) {
        if constexpr (sizeof(Real) == 4) {
            const int sl = slot<M>(idx);
            q[sl] = Real(s_u[DENS][sl] * s_p64[sl]);
        } else {
            const int sl = slot<M>(idx);
            double uu[5];
#pragma unroll
            for (int h = 0; h < 5; ++h) uu[h] = s_u[h][sl];
            q[sl] = uu[DENS] * pressure(a.gas64, uu);
        }
    }
    __syncthreads();
    const int el = tid / M2, line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const int base = el * M3;
    {  // x: t1(a,j,k) = sum_i Vinv(a,i) q(i,j,k)
        Real v[M];
#pragma unroll
        for (int i = 0; i < M; ++i) v[i] = q[base + swz<M>(i + M * ta + M2 * tb)];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            Real s = Real(0);
#pragma unroll
            for (int i = 0; i < M; ++i) s += a.op.Vinv[r][i] * v[i];
            t1[base + swz<M>(r + M * ta + M2 * tb)] = s;
        }
    }
    __syncthreads();
    {  // y: t2(i,b,k) = sum_j Vinv(b,j) t1(i,j,k)
        Real v[M];
#pragma unroll
        for (int j = 0; j < M; ++j) v[j] = t1[base + swz<M>(ta + M * j + M2 * tb)];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            Real s = Real(0);
#pragma unroll
            for (int j = 0; j < M; ++j) s += a.op.Vinv[r][j] * v[j];
            t2[base + swz<M>(ta + M * r + M2 * tb)] = s;
        }
    }
    __syncthreads();
    {  // z: modal(a,b,c) = sum_k Vinv(c,k) t2(a,b,k); energies
        Real v[M];
#pragma unroll
        for (int k = 0; k < M; ++k) v[k] = t2[base + swz<M>(ta + M * tb + M2 * k)];
        Real tot = Real(0), c1 = Real(0), c2 = Real(0);
#pragma unroll
        for (int r = 0; r < M; ++r) {
            Real s = Real(0);
#pragma unroll
            for (int k = 0; k < M; ++k) s += a.op.Vinv[r][k] * v[k];
            const Real e2 = s * s;
            tot += e2;
            if (ta < M - 1 && tb < M - 1 && r < M - 1) c1 += e2;
            if (ta < M - 2 && tb < M - 2 && r < M - 2) c2 += e2;
        }
        // q (row 2) is dead after the x contraction; rows 2..4 take the partials
        s_acc[2][tid] = tot;
        s_acc[3][tid] = c1;
        s_acc[4][tid] = c2;
    }
    __syncthreads();
    if (line == 0) {
        Real tot = Real(0), c1 = Real(0), c2 = Real(0);
        for (int l = 0; l < M2; ++l) {
            tot += s_acc[2][el * M2 + l];
            c1 += s_acc[3][el * M2 + l];
            c2 += s_acc[4][el * M2 + l];
        }
        double energy = 0.0;
        if (tot > Real(0)) {
            const double f1 = double(tot - c1) / double(tot);
            const double f2 = double(tot - c2) / double(tot);
            energy = f1 > f2 ? f1 : f2;
        }
        double alpha = 1.0 / (1.0 + exp(-a.ind_s / a.ind_threshold * (energy - a.ind_threshold)));
        if (alpha < a.alpha_min)
            alpha = 0.0;
        else if (alpha > 1.0 - a.alpha_min)
            alpha = 1.0;
        if (alpha > a.alpha_max) alpha = a.alpha_max;
        s_alpha[el] = Real(alpha);
        const long long e = e0 + el;
        if (a.alpha_out && e < a.nelem) a.alpha_out[e] = alpha;
    }
    __syncthreads();
}

// Positivity repair of one element held in LDS (one thread; rare path):
// hierarchical 2x2x2 tile squeeze of kernel/mcfv/3d/kernel_validate_mod.f90:
// 159-274 with quadrature-weighted tile means (LGL nodes) so the element
// integral is kept.  Returns false if the element is still invalid.
template <int M, typename Real>
__device__ bool repair_element(const Gas<Real>& g, const Real* w, Real* const* su, int base) {
    auto node_ok = [&](int n) {
        Real uu[5];
        for (int h = 0; h < 5; ++h) uu[h] = su[h][base + swz<M>(n)];
        return isvalid(g, uu);
    };
    auto all_ok = [&]() {
        for (int n = 0; n < M * M * M; ++n)
            if (!node_ok(n)) return false;
        return true;
    };
    if (all_ok()) return true;
    const Real decr = Real(0.01);
    int nlevels = 0;
    while ((1 << nlevels) < M) ++nlevels;
    if ((1 << nlevels) != M) {
        Real means[5] = {0, 0, 0, 0, 0}, wsum = 0;
        for (int k = 0; k < M; ++k)
            for (int j = 0; j < M; ++j)
                for (int i = 0; i < M; ++i) {
                    const Real ww = w[i] * w[j] * w[k];
                    wsum += ww;
                    for (int h = 0; h < 5; ++h) means[h] += ww * su[h][base + swz<M>(i + M * j + M * M * k)];
                }
        for (int h = 0; h < 5; ++h) means[h] = means[h] / wsum;
        // squeeze all nodes towards the mean in repair_decr steps
        Real squeeze = Real(1);
        Real scale_prev = Real(1);
        while (squeeze > Real(0)) {
            squeeze = squeeze - decr;
            if (squeeze < Real(0)) squeeze = Real(0);
            const Real f = squeeze / scale_prev;  // rescale current deviations
            for (int n = 0; n < M * M * M; ++n)
                for (int h = 0; h < 5; ++h) {
                    Real* x = &su[h][base + swz<M>(n)];
                    *x = means[h] + (squeeze == Real(0) ? Real(0) : f) * (*x - means[h]);
                }
            scale_prev = squeeze;
            if (all_ok()) break;
        }
        return all_ok();
    }
    for (int level = 1; level <= nlevels; ++level) {
        const int mm = 1 << (level - 1);
        const int nn = 1 << (nlevels - level);
        for (int r = 0; r < mm; ++r)
            for (int p = 0; p < mm; ++p)
                for (int q = 0; q < mm; ++q) {
                    Real tiles[8][5], tw[8];
                    for (int t = 0; t < 8; ++t) {
                        tw[t] = Real(0);
                        for (int h = 0; h < 5; ++h) tiles[t][h] = Real(0);
                    }
                    for (int kk = 0; kk < 2; ++kk)
                        for (int jj = 0; jj < 2; ++jj)
                            for (int ii = 0; ii < 2; ++ii) {
                                const int t = ii + 2 * jj + 4 * kk;
                                for (int k = (2 * r + kk) * nn; k < (2 * r + kk + 1) * nn; ++k)
                                    for (int j = (2 * p + jj) * nn; j < (2 * p + jj + 1) * nn; ++j)
                                        for (int i = (2 * q + ii) * nn; i < (2 * q + ii + 1) * nn; ++i) {
                                            const Real ww = w[i] * w[j] * w[k];
                                            tw[t] += ww;
                                            for (int h = 0; h < 5; ++h)
                                                tiles[t][h] += ww * su[h][base + swz<M>(i + M * j + M * M * k)];
                                        }
                                for (int h = 0; h < 5; ++h) tiles[t][h] = tiles[t][h] / tw[t];
                            }
                    bool ok = true;
                    for (int t = 0; t < 8; ++t)
                        if (!isvalid(g, tiles[t])) ok = false;
                    if (ok) continue;
                    Real means[5], twsum = Real(0);
                    for (int t = 0; t < 8; ++t) twsum += tw[t];
                    for (int h = 0; h < 5; ++h) {
                        Real s = Real(0);
                        for (int t = 0; t < 8; ++t) s += tw[t] * tiles[t][h];
                        means[h] = s / twsum;
                    }
                    Real squeeze = Real(1);
                    while (squeeze > Real(0)) {
                        squeeze = squeeze - decr;
                        if (squeeze < Real(0)) squeeze = Real(0);
                        bool okt = true;
                        for (int t = 0; t < 8; ++t) {
                            Real tt[5];
                            for (int h = 0; h < 5; ++h) tt[h] = means[h] + squeeze * (tiles[t][h] - means[h]);
                            if (!isvalid(g, tt)) okt = false;
                        }
                        if (okt) break;
                    }
                    for (int k = 2 * r * nn; k < (2 * r + 2) * nn; ++k)
                        for (int j = 2 * p * nn; j < (2 * p + 2) * nn; ++j)
                            for (int i = 2 * q * nn; i < (2 * q + 2) * nn; ++i)
                                for (int h = 0; h < 5; ++h) {
                                    Real* x = &su[h][base + swz<M>(i + M * j + M * M * k)];
                                    *x = means[h] + squeeze * (*x - means[h]);
                                }
                }
    }
    return all_ok();
}

// CFL quantities of one block of elements held in LDS (kernel/dgsem/weakform/
// 2d/kernel_timestep_mod.f90:65-121 with the 3-D minimum of kernel/mcfv/3d/
// kernel_timestep_mod.f90:60-106): per element dt_e = min_d delta_d /
// max_nodes lambda_d, the fluid speed maximum for GLM cleaning, reduced to
// one partial (dt_min, lambda_max, v_max) per block.  Thread t first folds
// the M nodes of x-line t, then one thread per element folds its lines.
template <int M, int BE, int R>
__device__ __forceinline__ void block_cfl(const Gas<double>& g, const double* __restrict__ delta, long long e0,
                                          long long nelem, double (*s_u)[BE * M * M * M], double* s_wbase,
                                          double* __restrict__ out) {
    static_assert(R >= BE * M * M, "work rows must hold one value per thread");
    double (*s_w)[R] = reinterpret_cast<double (*)[R]>(s_wbase);
    constexpr int M2 = M * M, M3 = M2 * M;
    const int tid = threadIdx.x;
    const int el = tid / M2, line = tid - el * M2;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, vm = 0.0;
    if (e0 + el < nelem) {
        for (int i = 0; i < M; ++i) {
            const int n = el * M3 + swz<M>(i + M * line);
            double uu[5];
#pragma unroll
            for (int h = 0; h < 5; ++h) uu[h] = s_u[h][n];
            const double a0 = lambda_max<0>(g, uu), a1 = lambda_max<1>(g, uu), a2 = lambda_max<2>(g, uu);
            l0 = a0 > l0 ? a0 : l0;
            l1 = a1 > l1 ? a1 : l1;
            l2 = a2 > l2 ? a2 : l2;
            const double v0 = fabs(uu[1] / uu[0]), v1 = fabs(uu[2] / uu[0]), v2 = fabs(uu[3] / uu[0]);
            double v = v0 > v1 ? v0 : v1;
            v = v > v2 ? v : v2;
            vm = v > vm ? v : vm;
        }
    }
    __syncthreads();  // the work rows may still be read before this point
    s_w[0][tid] = l0;
    s_w[1][tid] = l1;
    s_w[2][tid] = l2;
    s_w[3][tid] = vm;
    __syncthreads();
    double dte = 1.7976931348623157e308, lme = 0.0, vme = 0.0;
    if (tid < BE && e0 + tid < nelem) {
        const long long e = e0 + tid;
        double m0 = 0.0, m1 = 0.0, m2 = 0.0;
        for (int l = 0; l < M2; ++l) {
            const int r = tid * M2 + l;
            m0 = s_w[0][r] > m0 ? s_w[0][r] : m0;
            m1 = s_w[1][r] > m1 ? s_w[1][r] : m1;
            m2 = s_w[2][r] > m2 ? s_w[2][r] : m2;
            vme = s_w[3][r] > vme ? s_w[3][r] : vme;
        }
        const double t0 = delta[e * 3 + 0] / m0, t1 = delta[e * 3 + 1] / m1, t2 = delta[e * 3 + 2] / m2;
        dte = t0 < t1 ? t0 : t1;
        dte = dte < t2 ? dte : t2;
        lme = m0 > m1 ? m0 : m1;
        lme = lme > m2 ? lme : m2;
    }
    __syncthreads();
    if (tid < BE) {
        s_w[0][tid] = dte;
        s_w[1][tid] = lme;
        s_w[2][tid] = vme;
    }
    __syncthreads();
    if (tid == 0) {
        double dtb = 1.7976931348623157e308, lmb = 0.0, vmb = 0.0;
        for (int k = 0; k < BE; ++k) {
            dtb = s_w[0][k] < dtb ? s_w[0][k] : dtb;
            lmb = s_w[1][k] > lmb ? s_w[1][k] : lmb;
            vmb = s_w[2][k] > vmb ? s_w[2][k] : vmb;
        }
        out[0] = dtb;
        out[1] = lmb;
        out[2] = vmb;
    }
}

template <int M, int BE, int VOL, int SURF, typename Real>
__global__ __launch_bounds__(BE * M * M) void stage_kernel(StageArgs<M, Real> a) {
    using Sh = StageShape<M, BE>;
    constexpr int M3 = Sh::M3, NN = Sh::NN, NT = Sh::NT;
    constexpr bool F32 = sizeof(Real) == 4;
    __shared__ double s_u[5][NN];          // FP64 state: RK update, repair, CFL
    __shared__ Real s_acc[5][NN];          // RHS accumulators in operator precision
    __shared__ double s_p64[F32 ? NN : 1]; // FP32 operator: FP64 pressure per node
    __shared__ Real s_alpha[BE];
    __shared__ int s_bad[BE];

    const int tid = threadIdx.x;
    const long long nblocks = (a.nelem + BE - 1) / BE;
    const long long blk = xcd_block(nblocks);
    const long long e0 = blk * BE;
    const long long EM3 = a.nelem * (long long)M3;
    const long long nvalid_nodes = (a.nelem - e0 < BE ? a.nelem - e0 : BE) * (long long)M3;

    // ---- stage the block's state in LDS (coalesced) --------------------
#pragma unroll
    for (int h = 0; h < 5; ++h)
        for (int idx = tid; idx < NN; idx += NT)
            s_u[h][slot<M>(idx)] =
                idx < nvalid_nodes ? a.u_in[h * EM3 + e0 * M3 + idx] : (h == DENS || h == ENER ? 1.0 : 0.0);
    if (tid < BE) {
        s_alpha[tid] = Real(0);
        s_bad[tid] = 0;
    }
    __syncthreads();
    if constexpr (F32) {
        for (int idx = tid; idx < NN; idx += NT) {
            const int sl = slot<M>(idx);
            double uu[5];
#pragma unroll
            for (int h = 0; h < 5; ++h) uu[h] = s_u[h][sl];
            s_p64[sl] = pressure(a.gas64, uu);
        }
        __syncthreads();
    }

    if (a.indicator == IND_PP) indicator_pp<M, BE, Real>(a, e0, s_u, s_acc, s_alpha, s_p64);

    // ---- three direction passes ----------------------------------------
    direction_pass<0, M, BE, VOL, SURF, Real>(a, e0, s_u, s_acc, s_alpha, s_p64);
    __syncthreads();
    direction_pass<1, M, BE, VOL, SURF, Real>(a, e0, s_u, s_acc, s_alpha, s_p64);
    __syncthreads();
    direction_pass<2, M, BE, VOL, SURF, Real>(a, e0, s_u, s_acc, s_alpha, s_p64);
    __syncthreads();

    // ---- RHS only (parity path) -----------------------------------------
    if (a.rhs_out) {
#pragma unroll
        for (int h = 0; h < 5; ++h)
            for (int idx = tid; idx < nvalid_nodes; idx += NT)
                a.rhs_out[h * EM3 + e0 * M3 + idx] = double(s_acc[h][slot<M>(idx)]);
        return;
    }

    // ---- RK probe: LSRK 5-4 (low-storage/4-5/timedisc_probe_mod.f90:31-37)
    //      or SSPRK(5,4) (ssp/4-5/timedisc_probe_mod.f90:24-46) ---------------
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const int s = a.stage;
#pragma unroll
    for (int h = 0; h < 5; ++h)
        for (int idx = tid; idx < nvalid_nodes; idx += NT) {
            const long long gi = h * EM3 + e0 * M3 + idx;
            const int sl = slot<M>(idx);
            const double rhs = double(s_acc[h][sl]);
            const double u = s_u[h][sl];
            double st;
            if (a.timedisc == TD_LSRK54) {
                const double reg = (s == 1) ? rhs : a.rk.lsrk_a[s - 1] * a.r0[gi] + rhs;
                st = u + a.rk.lsrk_c[s - 1] * dt * reg;
                a.r0[gi] = reg;
            } else {
                const double* k = a.rk.ssp;
                switch (s) {
                case 1: st = u + k[3] * dt * rhs; a.r0[gi] = u; break;
                case 2: st = k[5] * a.r0[gi] + k[6] * u + k[7] * dt * rhs; break;
                case 3: st = k[9] * a.r0[gi] + k[10] * u + k[11] * dt * rhs; break;
                case 4:
                    st = k[13] * a.r0[gi] + k[14] * u + k[15] * dt * rhs;
                    a.r1[gi] = a.r1[gi] + k[22] * dt * rhs;
                    break;
                default: st = k[17] * a.r0[gi] + k[18] * u + k[19] * dt * rhs + a.r1[gi]; break;
                }
            }
            s_u[h][sl] = st;
        }
    __syncthreads();

    // ---- positivity repair (kernel_repair after the probe) ---------------
    if (a.repair) {
        for (int idx = tid; idx < nvalid_nodes; idx += NT) {
            double uu[5];
#pragma unroll
            for (int h = 0; h < 5; ++h) uu[h] = s_u[h][slot<M>(idx)];
            if (!isvalid(a.gas64, uu)) s_bad[idx / M3] = 1;
        }
        __syncthreads();
        if (tid < BE && s_bad[tid]) {
            double* su[5] = {s_u[0], s_u[1], s_u[2], s_u[3], s_u[4]};
            if (!repair_element<M, double>(a.gas64, a.rk.w64, su, tid * M3)) atomicOr(a.status, 1);
        }
        __syncthreads();
    }

    // ---- write the new state; SSP register r1 takes the repaired state ---
#pragma unroll
    for (int h = 0; h < 5; ++h)
        for (int idx = tid; idx < nvalid_nodes; idx += NT) {
            const long long gi = h * EM3 + e0 * M3 + idx;
            const double st = s_u[h][slot<M>(idx)];
            a.u_out[gi] = st;
            if (a.timedisc == TD_SSPRK54) {
                if (s == 2) a.r1[gi] = a.rk.ssp[20] * st;
                else if (s == 3) a.r1[gi] = a.r1[gi] + a.rk.ssp[21] * st;
            }
        }

    // ---- CFL partials of the new state (last stage of a step) -----------
    if (a.cfl_part)
        block_cfl<M, BE, F32 ? NN / 2 : NN>(a.gas64, a.delta, e0, a.nelem, s_u, reinterpret_cast<double*>(&s_acc[0][0]),
                                          a.cfl_part + 4 * blk);
}

// Standalone CFL reduction of a state in HBM (first step, calc_timestep):
// the same per-block partials as the fused stage epilogue.
template <int M, int BE>
__global__ __launch_bounds__(BE * M * M) void cfl_kernel(const double* __restrict__ u, const double* __restrict__ delta,
                                                         long long nelem, Gas<double> g, double* __restrict__ part) {
    constexpr int M3 = M * M * M, NN = BE * M3, NT = BE * M * M;
    __shared__ double s_u[5][NN];
    __shared__ double s_w[4][NT];
    const int tid = threadIdx.x;
    const long long blk = blockIdx.x;
    const long long e0 = blk * BE;
    const long long EM3 = nelem * (long long)M3;
    const long long nvalid = (nelem - e0 < BE ? nelem - e0 : BE) * (long long)M3;
#pragma unroll
    for (int h = 0; h < 5; ++h)
        for (int idx = tid; idx < NN; idx += NT)
            s_u[h][slot<M>(idx)] = idx < nvalid ? u[h * EM3 + e0 * M3 + idx] : (h == 0 || h == 4 ? 1.0 : 0.0);
    __syncthreads();
    block_cfl<M, BE, NT>(g, delta, e0, nelem, s_u, &s_w[0][0], part + 4 * blk);
}

// Deterministic reduction of the per-block partials (min and max are exact,
// so the result does not depend on the order) and
// dt = 0.5*CFL/(2*N_NODES-1) * dtmin (kernel_timestep_mod.f90:49).
__global__ __launch_bounds__(1024) void cfl_finalize(const double* __restrict__ part, long long npart, double cfl,
                                                      int M, double* __restrict__ dt_dev, double* __restrict__ hist,
                                                      long long slot) {
    __shared__ double s[3][1024];
    const int tid = threadIdx.x;
    double dt = 1.7976931348623157e308, lm = 0.0, vm = 0.0;
    for (long long i = tid; i < npart; i += 1024) {
        const double* q = part + 4 * i;
        dt = q[0] < dt ? q[0] : dt;
        lm = q[1] > lm ? q[1] : lm;
        vm = q[2] > vm ? q[2] : vm;
    }
    s[0][tid] = dt;
    s[1][tid] = lm;
    s[2][tid] = vm;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (tid < off) {
            s[0][tid] = s[0][tid + off] < s[0][tid] ? s[0][tid + off] : s[0][tid];
            s[1][tid] = s[1][tid + off] > s[1][tid] ? s[1][tid + off] : s[1][tid];
            s[2][tid] = s[2][tid + off] > s[2][tid] ? s[2][tid + off] : s[2][tid];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double d = 0.5 * cfl / double(2 * M - 1) * s[0][0];
        dt_dev[0] = d;
        dt_dev[1] = s[1][0];
        dt_dev[2] = s[2][0];
        if (hist && slot >= 0) hist[slot] = d;
    }
}

}  // namespace nemo
