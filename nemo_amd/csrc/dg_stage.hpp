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

#include <type_traits>

#include "dg_common.hpp"
#include "dg_euler.hpp"
#include "dg_mhd.hpp"
#include "mlp.hpp"

namespace nemo {

// Modal energies of a nodal field q held in LDS (block slot order) in the
// normalised-Legendre basis (utils/interpol_mod.f90:274-311, 584-601): the
// sum-factorised transform x -> y -> z, after which the calling thread holds
// its line's partial sums of the total energy (tot), of the modes with all
// indices below M-1 (c1), of the top shell (largest index M-1: s1) and of
// the second shell (largest index M-2: sh2).  Every sum is direct, free of
// the cancellation in tot - c1 or c1 - c2 when the high modes are small.
// q must be complete (barrier) on entry; t1 and t2 are work rows; t1 is
// free again on return, t2 only after a barrier.
template <int M, int BE, typename Real>
__device__ __forceinline__ void modal_partials(const StageArgs<M, Real>& a, const Real* q, Real* t1, Real* t2,
                                               Real& tot, Real& c1, Real& s1, Real& sh2) {
    constexpr int M2 = M * M, M3 = M2 * M;
    const int tid = threadIdx.x;
    const int el = tid / M2, line = tid - el * M2;
    const int ta = line % M, tb = line / M;
    const int base = el * M3;
    {  // x: t1(a,j,k) = sum_i Vinv(a,i) q(i,j,k)
        Real v[M];
#pragma unroll
        for (int i = 0; i < M; ++i) v[i] = q[bslot<M>(base, i + M * ta + M2 * tb)];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            Real s = Real(0);
#pragma unroll
            for (int i = 0; i < M; ++i) s += a.op.Vinv[r][i] * v[i];
            t1[bslot<M>(base, r + M * ta + M2 * tb)] = s;
        }
    }
    __syncthreads();
    {  // y: t2(i,b,k) = sum_j Vinv(b,j) t1(i,j,k)
        Real v[M];
#pragma unroll
        for (int j = 0; j < M; ++j) v[j] = t1[bslot<M>(base, ta + M * j + M2 * tb)];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            Real s = Real(0);
#pragma unroll
            for (int j = 0; j < M; ++j) s += a.op.Vinv[r][j] * v[j];
            t2[bslot<M>(base, ta + M * r + M2 * tb)] = s;
        }
    }
    __syncthreads();
    {  // z: modal(a,b,c) = sum_k Vinv(c,k) t2(a,b,k); energies
        Real v[M];
#pragma unroll
        for (int k = 0; k < M; ++k) v[k] = t2[bslot<M>(base, ta + M * tb + M2 * k)];
        tot = c1 = s1 = sh2 = Real(0);
#pragma unroll
        for (int r = 0; r < M; ++r) {
            Real s = Real(0);
#pragma unroll
            for (int k = 0; k < M; ++k) s += a.op.Vinv[r][k] * v[k];
            const Real e2 = s * s;
            const int top = (ta > tb ? ta : tb) > r ? (ta > tb ? ta : tb) : r;
            tot += e2;
            if (top == M - 1)
                s1 += e2;
            else
                c1 += e2;
            if (top == M - 2) sh2 += e2;
        }
    }
}

// Blending factor from the modal energy of Hennemann et al. (JCP 426, 2021,
// sec. 3): max(top shell / total, second shell / energy below the top
// shell); logistic ramp with sharpness s = log(9999) around the threshold
// 0.5*10^(-1.8*M^0.25), values below alpha_min set to 0 (above 1-alpha_min
// to 1), clipped at alpha_max.
template <int M, typename Real>
__device__ __forceinline__ double pp_alpha(const StageArgs<M, Real>& a, Real tot, Real s1, Real c1, Real sh2) {
    double energy = 0.0;
    if (tot > Real(0)) {
        const double f1 = double(s1) / double(tot);
        const double f2 = c1 > Real(0) ? double(sh2) / double(c1) : 0.0;
        energy = f1 > f2 ? f1 : f2;
    }
    double alpha = 1.0 / (1.0 + exp(-a.ind_s / a.ind_threshold * (energy - a.ind_threshold)));
    if (alpha < a.alpha_min)
        alpha = 0.0;
    else if (alpha > 1.0 - a.alpha_min)
        alpha = 1.0;
    if (alpha > a.alpha_max) alpha = a.alpha_max;
    return alpha;
}

// Persson-Peraire indicator on rho*p (Hennemann et al. 2021, sec. 3).
// Returns the blending factor of the calling thread's element (one
// workgroup is one wavefront, so it is broadcast by a lane shuffle); with
// write_out the factor also goes to alpha_out.
template <int EQ, int M, int BE, typename Real, typename TU>
__device__ __forceinline__ Real indicator_pp(const StageArgs<M, Real>& a, long long e0,
                                             const TU (*s_u)[BE * M * M * M], Real (*s_acc)[BE * M * M * M],
                                             const double* s_p64, bool write_out = true) {
    constexpr int V = NVars<EQ>::V;
    using Sh = StageShape<M, BE>;
    constexpr int M2 = Sh::M2, NN = Sh::NN, NT = Sh::NT;
    const int tid = threadIdx.x;
    Real* q = s_acc[2];
    for (int idx = tid; idx < NN; idx += NT) {
        const int sl = slot<M>(idx);
        if constexpr (std::is_same<TU, float>::value) {
            q[sl] = Real(double(s_u[DENS][sl]) * s_p64[sl]);
        } else {
            double uu[V];
#pragma unroll
            for (int h = 0; h < V; ++h) uu[h] = s_u[h][sl];
            q[sl] = Real(uu[DENS] * pressure64<EQ>(a.ph64, uu));
        }
    }
    __syncthreads();
    Real tot, c1, s1, sh2;
    modal_partials<M, BE, Real>(a, q, s_acc[0], s_acc[1], tot, c1, s1, sh2);
    // rows 0 (t1) and 2 (q) are dead after the y contraction; rows 0, 2..4
    // take the partials
    s_acc[0][tid] = tot;
    s_acc[2][tid] = s1;
    s_acc[3][tid] = c1;
    s_acc[4][tid] = sh2;
    __syncthreads();
    const int el = tid / M2, line = tid - el * M2;
    Real alpha_el = Real(0);
    if (line == 0) {
        Real st = Real(0), t1 = Real(0), tc = Real(0), t2 = Real(0);
        for (int l = 0; l < M2; ++l) {
            st += s_acc[0][el * M2 + l];
            t1 += s_acc[2][el * M2 + l];
            tc += s_acc[3][el * M2 + l];
            t2 += s_acc[4][el * M2 + l];
        }
        const double alpha = pp_alpha(a, st, t1, tc, t2);
        alpha_el = Real(alpha);
        const long long e = e0 + el;
        if (write_out && a.alpha_out && e < a.nelem) a.alpha_out[e] = alpha;
    }
    alpha_el = __shfl(alpha_el, el * M2);
    __syncthreads();  // the work rows are the direction passes' accumulators
    return alpha_el;
}

// Element sums of the line partials over the 16 lanes of an element (M = 4).
template <typename Real>
__device__ __forceinline__ Real sum16(Real v) {
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Counter-based sampling key of (element, step).
__device__ __forceinline__ uint32_t sample_hash(long long e, uint32_t step) {
    uint32_t h = uint32_t(e) * 0x9E3779B1u ^ (uint32_t(e >> 32) * 0x85EBCA77u) ^ (step * 0xC2B2AE3Du);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

// Learned indicator (M = 4, one element per 16 lanes).  At the first stage
// of a step (alpha_refresh) it computes per element the features
//   0.1 ln(f + 1e-12),  f = E(top shell)/E, E(top two shells)/E of rho and p,
// evaluates the network, and lets the Persson-Peraire indicator judge a
// sampled fraction of the elements (sample_hash < sample_thresh): where the
// prediction falls short of the PP factor by more than judge_tol, or where
// any node has rho <= 0 or p <= 0, the PP factor is used.  The factor is
// kept in alpha_step for the remaining stages of the step.  With feat_out
// (training data) it writes the features and the PP factor of every element.
template <int EQ, int M, int BE, typename Real, typename TU>
__device__ __forceinline__ Real indicator_mlp(const StageArgs<M, Real>& a, long long e0,
                                              const TU (*s_u)[BE * M * M * M], Real (*s_acc)[BE * M * M * M],
                                              const double* s_p64) {
    static_assert(M == 4 && BE * M * M == 64, "the learned indicator maps 16 lanes to an element");
    constexpr int V = NVars<EQ>::V;
    constexpr int M2 = 16, NN = BE * 64, NT = 64;
    const int tid = threadIdx.x;
    const int el = tid / M2;
    const long long e = e0 + el;
    const bool valid = e < a.nelem;
    if (!a.alpha_refresh) return valid ? Real(a.alpha_step[e]) : Real(0);

    Real* q = s_acc[2];
    // rho, then p; positivity flags per element
    uint32_t bad = 0;
    float fr1, fr2, fp1, fp2;
    for (int pass = 0; pass < 2; ++pass) {
        for (int idx = tid; idx < NN; idx += NT) {
            const int sl = slot<M>(idx);
            double rho, p;
            if constexpr (std::is_same<TU, float>::value) {
                rho = double(s_u[DENS][sl]);
                p = s_p64[sl];
            } else {
                double uu[V];
#pragma unroll
                for (int h = 0; h < V; ++h) uu[h] = s_u[h][sl];
                rho = uu[DENS];
                p = pressure64<EQ>(a.ph64, uu);
            }
            q[sl] = Real(pass == 0 ? rho : p);
            if (pass == 0 && !(rho > 0.0 && p > 0.0) && e0 + idx / 64 < a.nelem) bad |= 1u << (idx / 64);
        }
        __syncthreads();
        Real tot, c1, s1, sh2;
        modal_partials<M, BE, Real>(a, q, s_acc[0], s_acc[1], tot, c1, s1, sh2);
        Real s2 = s1 + sh2;  // top two shells
        tot = sum16(tot);
        s1 = sum16(s1);
        s2 = sum16(s2);
        const float f1 = tot > Real(0) ? float(s1 / tot) : 0.0f;
        const float f2 = tot > Real(0) ? float(s2 / tot) : 0.0f;
        if (pass == 0) {
            fr1 = f1;
            fr2 = f2;
        } else {
            fp1 = f1;
            fp2 = f2;
        }
        __syncthreads();
    }
    const float x[4] = {mlp_feature(fr1), mlp_feature(fr2), mlp_feature(fp1), mlp_feature(fp2)};
    uint32_t bad_el = 0;
#pragma unroll
    for (int k = 0; k < BE; ++k)
        if (__any((bad >> k) & 1u)) bad_el |= 1u << k;

    // judge / fallback / training data need the PP factor
    // judged blocks are drawn whole (key: the block's first element), so
    // that the Persson-Peraire factor is computed for the sampled fraction
    // of the wavefronts rather than for every wavefront holding one sampled
    // element
    const bool judged = valid && (a.feat_out || sample_hash(e0, a.step_index) < a.sample_thresh);
    const bool need_pp = __any(judged) || bad_el != 0;
    Real a_pp = Real(0);
    if (need_pp) a_pp = indicator_pp<EQ, M, BE, Real, TU>(a, e0, s_u, s_acc, s_p64, false);

    if (a.feat_out) {  // training data: features and PP factor of every element
        if ((tid & 15) == 0 && valid) {
#pragma unroll
            for (int k = 0; k < 4; ++k) a.feat_out[e * 4 + k] = x[k];
            if (a.alpha_out) a.alpha_out[e] = double(a_pp);
        }
        return a_pp;
    }

    float feat[BE][4], y[BE];
#pragma unroll
    for (int k = 0; k < BE; ++k)
#pragma unroll
        for (int f = 0; f < 4; ++f) feat[k][f] = lane_bcast(x[f], k * 16);
    mlp_forward_wave<BE>(a.mlp, feat, y);
    float ym = y[0];
#pragma unroll
    for (int k = 1; k < BE; ++k)
        if (el == k) ym = y[k];
    double alpha = double(ym);
    if (alpha < a.alpha_min)
        alpha = 0.0;
    else if (alpha > 1.0 - a.alpha_min)
        alpha = 1.0;
    if (alpha > a.alpha_max) alpha = a.alpha_max;
    int code = 0;  // 1 judged, 2 judged and overridden, 4 fallback
    if (judged) {
        code = 1;
        if (double(a_pp) > alpha + a.judge_tol) {
            alpha = double(a_pp);
            code = 3;
        }
    }
    if ((bad_el >> el) & 1u) {
        alpha = double(a_pp);
        code |= 4;
    }
    // A diagnostic launch (RHS, indicator or feature output, rhs_out set)
    // reports the factor through alpha_out only: the step's factor and the
    // judge's counters belong to the stage sequence.
    if ((tid & 15) == 0 && valid) {
        if (!a.rhs_out) a.alpha_step[e] = alpha;
        if (a.alpha_out) a.alpha_out[e] = alpha;
        if (a.mlp_stats && code && !a.rhs_out) {
            if (code & 1) atomicAdd(&a.mlp_stats[0], 1ull);
            if (code & 2) atomicAdd(&a.mlp_stats[1], 1ull);
            if (code & 4) atomicAdd(&a.mlp_stats[2], 1ull);
        }
    }
    return Real(alpha);
}

// Positivity repair of one element held in LDS (one thread; rare path):
// hierarchical 2x2x2 tile squeeze of kernel/mcfv/3d/kernel_validate_mod.f90:
// 159-274 with quadrature-weighted tile means (LGL nodes) so the element
// integral is kept.  Returns false if the element is still invalid.
template <int EQ, int M>
__device__ bool repair_element(const Phys64& P, const double* w, double* const* su, int base) {
    using Real = double;
    constexpr int V = NVars<EQ>::V;
    auto node_ok = [&](int n) {
        Real uu[V];
        for (int h = 0; h < V; ++h) uu[h] = su[h][bslot<M>(base, n)];
        return isvalid64<EQ>(P, uu);
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
        Real means[V] = {}, wsum = 0;
        for (int k = 0; k < M; ++k)
            for (int j = 0; j < M; ++j)
                for (int i = 0; i < M; ++i) {
                    const Real ww = w[i] * w[j] * w[k];
                    wsum += ww;
                    for (int h = 0; h < V; ++h) means[h] += ww * su[h][bslot<M>(base, i + M * j + M * M * k)];
                }
        for (int h = 0; h < V; ++h) means[h] = means[h] / wsum;
        // squeeze all nodes towards the mean in repair_decr steps
        Real squeeze = Real(1);
        Real scale_prev = Real(1);
        while (squeeze > Real(0)) {
            squeeze = squeeze - decr;
            if (squeeze < Real(0)) squeeze = Real(0);
            const Real f = squeeze / scale_prev;  // rescale current deviations
            for (int n = 0; n < M * M * M; ++n)
                for (int h = 0; h < V; ++h) {
                    Real* x = &su[h][bslot<M>(base, n)];
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
                    Real tiles[8][V], tw[8];
                    for (int t = 0; t < 8; ++t) {
                        tw[t] = Real(0);
                        for (int h = 0; h < V; ++h) tiles[t][h] = Real(0);
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
                                            for (int h = 0; h < V; ++h)
                                                tiles[t][h] += ww * su[h][bslot<M>(base, i + M * j + M * M * k)];
                                        }
                                for (int h = 0; h < V; ++h) tiles[t][h] = tiles[t][h] / tw[t];
                            }
                    bool ok = true;
                    for (int t = 0; t < 8; ++t)
                        if (!isvalid64<EQ>(P, tiles[t])) ok = false;
                    if (ok) continue;
                    Real means[V], twsum = Real(0);
                    for (int t = 0; t < 8; ++t) twsum += tw[t];
                    for (int h = 0; h < V; ++h) {
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
                            Real tt[V];
                            for (int h = 0; h < V; ++h) tt[h] = means[h] + squeeze * (tiles[t][h] - means[h]);
                            if (!isvalid64<EQ>(P, tt)) okt = false;
                        }
                        if (okt) break;
                    }
                    for (int k = 2 * r * nn; k < (2 * r + 2) * nn; ++k)
                        for (int j = 2 * p * nn; j < (2 * p + 2) * nn; ++j)
                            for (int i = 2 * q * nn; i < (2 * q + 2) * nn; ++i)
                                for (int h = 0; h < V; ++h) {
                                    Real* x = &su[h][bslot<M>(base, i + M * j + M * M * k)];
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
template <int EQ, int M, int BE, int R>
__device__ __forceinline__ void block_cfl(const Phys64& P, const double* __restrict__ delta, long long e0,
                                          long long nelem, const double (*s_u)[BE * M * M * M], double* s_wbase,
                                          double* __restrict__ out) {
    static_assert(R >= BE * M * M, "work rows must hold one value per thread");
    double (*s_w)[R] = reinterpret_cast<double (*)[R]>(s_wbase);
    constexpr int M2 = M * M, M3 = M2 * M;
    const int tid = threadIdx.x;
    const int el = tid / M2, line = tid - el * M2;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, vm = 0.0;
    if (e0 + el < nelem) {
        for (int i = 0; i < M; ++i) {
            const int n = bslot<M>(el * M3, i + M * line);
            double uu[NVars<EQ>::V];
#pragma unroll
            for (int h = 0; h < NVars<EQ>::V; ++h) uu[h] = s_u[h][n];
            const double a0 = lambda_max64<EQ, 0>(P, uu), a1 = lambda_max64<EQ, 1>(P, uu),
                         a2 = lambda_max64<EQ, 2>(P, uu);
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

// LDS of one stage block.  FP32 operator: the operator reads an FP32 copy
// of the state (plus the FP64 pressure per node, for the indicator and the
// Euler HLLC/HLLE blend) and the FP64 new state of the RK update is laid
// over the operator's LDS afterwards, so a workgroup needs 12 KB instead of
// 17.5 KB (Euler, N=3) or 20 KB instead of 27 KB (GLM-MHD, N=3: eight
// workgroups per CU, two waves per SIMD, instead of five) and more of them
// fit on a CU.
template <int EQ, int M, int BE, typename Real>
struct StageLds {
    static constexpr int NN = BE * M * M * M, V = NVars<EQ>::V;
    static constexpr bool COMPACT = sizeof(Real) == 4;
    using TU = typename std::conditional<COMPACT, float, double>::type;
    static constexpr int OP_DOUBLES = (V * NN * int(sizeof(TU)) + V * NN * int(sizeof(Real)) + 7) / 8;
    static constexpr int RAW_DOUBLES = OP_DOUBLES > V * NN ? OP_DOUBLES : V * NN;
    static constexpr int DOUBLES = RAW_DOUBLES + (COMPACT ? NN : 0);
};

// LSRK 5-4 probe of one value (low-storage/4-5/timedisc_probe_mod.f90:31-37)
// with the register value r0 already loaded: reg = f (stage 1) or
// a_s r0 + f; u + c_s dt reg; r0 <- reg.
template <int M, typename Real>
__device__ __forceinline__ double probe_lsrk(const StageArgs<M, Real>& a, long long gi, int s, double dt, double u,
                                             double r0, double rhs) {
    const double reg = (s == 1) ? rhs : a.rk.lsrk_a[s - 1] * r0 + rhs;
    a.r0[gi] = reg;
    return u + a.rk.lsrk_c[s - 1] * dt * reg;
}

// One block of BE elements through a whole stage (the body of the stage
// kernels); lds holds StageLds::DOUBLES doubles.
template <int EQ, int M, int BE, int VOL, int SURF, typename Real, int IK, bool NEAR1_SPLIT = true>
__device__ __forceinline__ void stage_block(const StageArgs<M, Real>& a, const long long blk, double* lds) {
    using Sh = StageShape<M, BE>;
    using L = StageLds<EQ, M, BE, Real>;
    constexpr int M3 = Sh::M3, NN = Sh::NN, NT = Sh::NT;
    static_assert(NT <= 64, "a workgroup is one wavefront");
    constexpr int V = NVars<EQ>::V;
    constexpr bool F32 = sizeof(Real) == 4;
    constexpr bool COMPACT = L::COMPACT;
    using TU = typename L::TU;
    double* s_raw = lds;
    double* s_p64 = lds + L::RAW_DOUBLES;
    TU (*s_u)[NN] = reinterpret_cast<TU (*)[NN]>(&s_raw[0]);
    Real (*s_acc)[NN] = reinterpret_cast<Real (*)[NN]>(reinterpret_cast<TU*>(&s_raw[0]) + V * NN);
    double (*s_u64)[NN] = reinterpret_cast<double (*)[NN]>(&s_raw[0]);  // FP64 state (RK update, repair, CFL)

    const int tid = threadIdx.x;
    const long long e0 = blk * BE;
    const long long EM3 = a.nelem * (long long)M3;
    const long long nvalid_nodes = (a.nelem - e0 < BE ? a.nelem - e0 : BE) * (long long)M3;
    // GLM cleaning speed and damping of this step (equations_share_mod)
    double glm_alpha = 0.0;
    MhdGas<Real> mg = a.mhd;
    if constexpr (EQ == EQ_MHD) {
        mg.ch = Real(a.glm[0]);
        glm_alpha = a.glm[1];
    }

    // connectivity of the thread's element, requested with the state
    ElemGeom geo{};
    if constexpr (EQ == EQ_EULER) geo = load_geom(a, e0 + tid / Sh::M2);

    // ---- stage the block's state in LDS (coalesced) --------------------
    // Every thread owns Q = NN / NT = M node slots per variable; all its
    // loads are issued before the first is used, so the wavefront waits for
    // HBM once, not once per node and variable.
    static_assert(NN % NT == 0, "whole lines per thread");
    constexpr int QS = NN / NT;
    // FP64 two-point volume flux (Euler): whether the log-means may skip the
    // log's range reduction (NEAR1 of the direction passes), decided for the
    // whole block during staging so that each variant of the passes is
    // straight-line code.
    constexpr bool NEAR1_TEST = NEAR1_SPLIT && EQ == EQ_EULER && !F32 && VOL == VOL_EC;
    bool near1 = false;
    FaceNodes fn;
    {
        double uu[V][QS];
#pragma unroll
        for (int h = 0; h < V; ++h)
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                // Slots past the last element take node 0 of element 0 (any
                // valid state: their results are discarded).  An unconditional
                // load, not a select of one: the compiler would sink that into
                // a branch, after which every later load waits for it.
                const int idx = tid + q * NT;
                uu[h][q] = a.u_in[h * EM3 + (idx < nvalid_nodes ? e0 * M3 + idx : 0)];
            }
        if constexpr (EQ == EQ_EULER) {
            load_face_nodes<0, M, BE>(a, e0, geo, fn);
            load_face_nodes<1, M, BE>(a, e0, geo, fn);
            load_face_nodes<2, M, BE>(a, e0, geo, fn);
        }
#pragma unroll
        for (int q = 0; q < QS; ++q) {
            const int sl = slot<M>(tid + q * NT);
#pragma unroll
            for (int h = 0; h < V; ++h) s_u[h][sl] = TU(uu[h][q]);
            if constexpr (COMPACT) {
                double un[V];
#pragma unroll
                for (int h = 0; h < V; ++h) un[h] = uu[h][q];
                s_p64[sl] = pressure64<EQ>(a.ph64, un);
            }
        }
        if constexpr (NEAR1_TEST) {
            // Every node's density and pressure within a factor F of node 0's
            // (both positive): two densities of the block are then within
            // F^2 = 1.18 of each other and two betas rho/(2p) within F^4 =
            // 1.386 (1.4 with the few ulps by which the passes' pressures,
            // summed in their own order, differ from this one), so every
            // ratio the log-means take lies in [sqrt(1/2), sqrt(2)).  NaN
            // or non-positive states fail the test.
            constexpr double F = 1.085;
            const double rref = __shfl(uu[DENS][0], 0);
            const double eref = __shfl(uu[ENER][0], 0);
            const double kref = __shfl(uu[MOMX][0] * uu[MOMX][0] + uu[MOMY][0] * uu[MOMY][0] +
                                       uu[MOMZ][0] * uu[MOMZ][0], 0);
            const double km1 = a.ph64.gas.kappa_m1;
            const double pref = km1 * (eref - 0.5 * kref * rcp_nr(rref));
            bool ok = rref >= 1e-300 && pref >= 1e-300;
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                const double r = uu[DENS][q];
                const double k = uu[MOMX][q] * uu[MOMX][q] + uu[MOMY][q] * uu[MOMY][q] + uu[MOMZ][q] * uu[MOMZ][q];
                const double p = km1 * (uu[ENER][q] - 0.5 * k * rcp_nr(r));
                ok = ok && r <= F * rref && rref <= F * r && p <= F * pref && pref <= F * p;
            }
            // readfirstlane: a wavefront-uniform value the compiler can see
            // as such, so that the passes' two variants form a scalar branch
            // rather than two exec-masked regions (whose values would all be
            // live at once)
            near1 = __builtin_amdgcn_readfirstlane(int(__all(ok))) != 0;
        }
    }
    // FP64 Euler with the two-point volume flux, M <= 4, blocks that pass the
    // near-1 test: the LSRK register of the RK update is requested before
    // the direction passes (load_r0), so that it arrives during the operator
    // instead of being waited for at the end; its 40 VGPRs fit under the 256
    // of two waves per SIMD there (not in the general variant of the passes,
    // nor for the other volume forms or M = 5).  Unconditional loads --
    // stage 1 and the SSP scheme read u_in instead and ignore it -- keep the
    // wait counts exact.  The x pass then runs last and, with LSRK, applies
    // the update to its own line's nodes (XUPD), so the register is
    // requested in that layout: the M nodes of the thread's x line (QS = M).
    constexpr bool R0_EARLY = EQ == EQ_EULER && !F32 && VOL == VOL_EC && M <= 4;
    const bool xlast = R0_EARLY && a.timedisc == TD_LSRK54 && !a.rhs_out;
    auto load_r0 = [&](double (&r0e)[V][QS]) {
        static_assert(QS == M, "one x line per thread");
        const double* rsrc = (a.timedisc == TD_LSRK54 && a.stage != 1) ? a.r0 : a.u_in;
        const int lel = tid / Sh::M2, lline = tid - lel * Sh::M2;
        const long long ex = e0 + lel < a.nelem ? e0 + lel : 0;
#pragma unroll
        for (int h = 0; h < V; ++h)
#pragma unroll
            for (int q = 0; q < QS; ++q) {
                r0e[h][q] = rsrc[h * EM3 + ex * M3 + nidx<0, M>(q, lline % M, lline / M)];
            }
    };
    __syncthreads();

    Real alpha = Real(0);
    if constexpr (IK == IK_MLP) {
        alpha = indicator_mlp<EQ, M, BE, Real, TU>(a, e0, s_u, s_acc, s_p64);
    } else if constexpr (IK == IK_STORED) {
        const long long e = e0 + tid / Sh::M2;
        alpha = e < a.nelem ? Real(a.alpha_step[e]) : Real(0);
    } else if (a.indicator == IND_PP) {
        alpha = indicator_pp<EQ, M, BE, Real, TU>(a, e0, s_u, s_acc, s_p64);
    }

    // indicator only (nemo_gpu_indicator, Euler): the factor has gone to
    // alpha_out
    if constexpr (EQ == EQ_EULER)
        if (a.alpha_only) return;

    // ---- three direction passes ----------------------------------------
    if constexpr (EQ == EQ_MHD) {
        direction_pass_mhd<0, M, BE, VOL, SURF, Real>(a, mg, e0, s_u, s_acc, alpha);
        __syncthreads();
        direction_pass_mhd<1, M, BE, VOL, SURF, Real>(a, mg, e0, s_u, s_acc, alpha);
        __syncthreads();
        direction_pass_mhd<2, M, BE, VOL, SURF, Real>(a, mg, e0, s_u, s_acc, alpha);
    } else {
        surface_pass<0, M, BE, SURF, Real, TU>(a, e0, s_u, s_acc, s_p64, geo, fn);
        __syncthreads();
        surface_pass<1, M, BE, SURF, Real, TU>(a, e0, s_u, s_acc, s_p64, geo, fn);
        __syncthreads();
        surface_pass<2, M, BE, SURF, Real, TU>(a, e0, s_u, s_acc, s_p64, geo, fn);
        __syncthreads();
        if constexpr (R0_EARLY) {
            // x runs last (in the near-1 variant with LSRK it fuses the
            // update, otherwise it accumulates like the others)
            XUpdate<M> xu;
            xu.dt = a.dt_run ? a.dt_run[a.block_run[blk]] : (a.dt_dev ? *a.dt_dev : a.dt);
            // The near-1 variant fuses the update into the x pass and ends in
            // its own store of the new state (to HBM, or to LDS for the tail
            // below); the general variant (blocks with steep gradients)
            // accumulates x like the other passes and leaves the update to
            // the tail, so that no register is live across its passes.
            auto passes = [&](auto near1_tag) -> bool {
                constexpr bool NR = decltype(near1_tag)::value;
                // distinct empty markers keep the compiler from hoisting the
                // variants' common first instructions above the fork, where
                // their registers would be live in both variants (it spilled
                // over a hundred VGPRs without them)
                if constexpr (NR)
                    asm volatile("; near1 passes");
                else
                    asm volatile("; general passes");
                if constexpr (NR) load_r0(xu.r0);
                direction_pass<1, M, BE, VOL, SURF, Real, TU, false, NR>(a, e0, s_u, s_acc, alpha, s_p64, geo);
                __syncthreads();
                direction_pass<2, M, BE, VOL, SURF, Real, TU, false, NR>(a, e0, s_u, s_acc, alpha, s_p64, geo);
                __syncthreads();
                direction_pass<0, M, BE, VOL, SURF, Real, TU, NR, NR>(a, e0, s_u, s_acc, alpha, s_p64, geo, &xu,
                                                                      NR && xlast);
                if (!NR || !xlast) return false;
                // The new state of the thread's x line: stored straight to
                // HBM unless a repair or the CFL partials need the block in
                // LDS (positivity test as in the repair below).
                const int lel = tid / Sh::M2, lline = tid - lel * Sh::M2;
                const int lta = lline % M, ltb = lline / M;
                const bool lvalid = e0 + lel < a.nelem;
                bool direct = a.cfl_part == nullptr;
                if (a.repair) {
                    bool bad = false;
#pragma unroll
                    for (int q = 0; q < QS; ++q) {
                        double uu[V];
#pragma unroll
                        for (int h = 0; h < V; ++h) uu[h] = xu.unew[h][q];
                        if (!isvalid64<EQ>(a.ph64, uu)) bad = true;
                    }
                    if (__any(bad && lvalid)) direct = false;
                }
                if (direct) {
                    if (lvalid) {
                        const long long eb = (e0 + lel) * M3;
#pragma unroll
                        for (int h = 0; h < V; ++h)
                            if constexpr (M == 4) {  // 16-byte stores of consecutive node pairs
                                for (int q = 0; q < QS; q += 2)
                                    *reinterpret_cast<double2*>(a.u_out + h * EM3 + eb + nidx<0, M>(q, lta, ltb)) =
                                        make_double2(xu.unew[h][q], xu.unew[h][q + 1]);
                            } else {
                                for (int q = 0; q < QS; ++q) a.u_out[h * EM3 + eb + nidx<0, M>(q, lta, ltb)] = xu.unew[h][q];
                            }
                    }
                    return true;
                }
                __syncthreads();
#pragma unroll
                for (int h = 0; h < V; ++h)
#pragma unroll
                    for (int q = 0; q < QS; ++q) s_u64[h][bslot<M>(lel * M3, nidx<0, M>(q, lta, ltb))] = xu.unew[h][q];
                return false;
            };
            const bool done = near1 ? passes(BoolConst<true>{}) : passes(BoolConst<false>{});
            if (done) return;
        } else {
            auto passes = [&](auto near1_tag) {
                constexpr bool NR = decltype(near1_tag)::value;
                if constexpr (NR)
                    asm volatile("; near1 passes");
                else
                    asm volatile("; general passes");
                direction_pass<0, M, BE, VOL, SURF, Real, TU, false, NR>(a, e0, s_u, s_acc, alpha, s_p64, geo);
                __syncthreads();
                direction_pass<1, M, BE, VOL, SURF, Real, TU, false, NR>(a, e0, s_u, s_acc, alpha, s_p64, geo);
                __syncthreads();
                direction_pass<2, M, BE, VOL, SURF, Real, TU, false, NR>(a, e0, s_u, s_acc, alpha, s_p64, geo);
            };
            if (NEAR1_TEST && near1)
                passes(BoolConst<NEAR1_TEST>{});
            else
                passes(BoolConst<false>{});
        }
    }
    __syncthreads();

    // Time derivative of variable h at block node slot sl, in FP64; GLM
    // damping -alpha psi (kernel/mcfv/3d/kernel_flux_mod.f90:496-497) is part
    // of the FP64 divergence-cleaning update.  The compact layout holds psi
    // in FP32 only: its callers pass the node's FP64 value as u64.
    auto dudt = [&](int h, int sl, double u64) {
        double r = double(s_acc[h][sl]);
        if constexpr (EQ == EQ_MHD)
            if (h == GLMP) r = r - glm_alpha * (COMPACT ? u64 : double(s_u[GLMP][sl]));
        return r;
    };

    // ---- RHS only (parity path) -----------------------------------------
    if (a.rhs_out) {
#pragma unroll
        for (int h = 0; h < V; ++h)
            for (int idx = tid; idx < nvalid_nodes; idx += NT)
                a.rhs_out[h * EM3 + e0 * M3 + idx] =
                    dudt(h, slot<M>(idx), COMPACT ? a.u_in[h * EM3 + e0 * M3 + idx] : 0.0);
        return;
    }

    // ---- RK probe: LSRK 5-4 (low-storage/4-5/timedisc_probe_mod.f90:31-37)
    //      or SSPRK(5,4) (ssp/4-5/timedisc_probe_mod.f90:24-46) ---------------
    const double dt = a.dt_run ? a.dt_run[a.block_run[blk]] : (a.dt_dev ? *a.dt_dev : a.dt);
    const int s = a.stage;
    auto probe = [&](long long gi, double u, double rhs) {
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
        return st;
    };
    // Each thread owns Q node slots per variable.  The loads of a variable
    // (registers, and the FP64 state for the compact layout) are issued
    // together before any is used, so a wavefront waits for memory once per
    // variable instead of once per node.
    constexpr int Q = (NN + NT - 1) / NT;
    const bool lsrk = a.timedisc == TD_LSRK54;
    if constexpr (COMPACT) {
        // FP64 state from HBM, new state in registers until the operator's
        // LDS is free.  The register and the new state stream through L2
        // as nontemporal accesses, which it evicts first, so that more of
        // the state the update re-reads here is still in L2 (FP32 + PP
        // stage 3.6 % faster; the FP64 kernel, whose fused x pass stores
        // 32-byte pieces per instruction, ran 11 % slower with them).
        // With LSRK the FP32 operator keeps its register in
        // FP32 (the r0 buffer read as floats, half the register traffic):
        // reg = a_s r0 + f is formed in FP64 and rounded once, and the state
        // advances by the rounded value, so the next stage continues from
        // the register it stored; u and its update stay FP64.
        float* const r0f = reinterpret_cast<float*>(a.r0);
        double st[V][Q];
#pragma unroll
        for (int h = 0; h < V; ++h) {
            double uin[Q], reg[Q];
            float reg32[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = tid + q * NT;
                const long long gi = h * EM3 + e0 * M3 + idx;
                const bool ok = idx < nvalid_nodes;
                uin[q] = ok ? a.u_in[gi] : 0.0;
                if (lsrk)
                    reg32[q] = (ok && s != 1) ? __builtin_nontemporal_load(&r0f[gi]) : 0.0f;
                else
                    reg[q] = ok ? a.r0[gi] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = tid + q * NT;
                st[h][q] = 0.0;
                if (idx < nvalid_nodes) {
                    const long long gi = h * EM3 + e0 * M3 + idx;
                    const double f = dudt(h, slot<M>(idx), uin[q]);
                    if (lsrk) {
                        // low-storage/4-5/timedisc_probe_mod.f90:31-37
                        const float rg = float(s == 1 ? f : a.rk.lsrk_a[s - 1] * double(reg32[q]) + f);
                        __builtin_nontemporal_store(rg, &r0f[gi]);
                        st[h][q] = uin[q] + a.rk.lsrk_c[s - 1] * dt * double(rg);
                    } else {
                        st[h][q] = probe(gi, uin[q], f);
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < V; ++h)
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = tid + q * NT;
                if (idx < nvalid_nodes) s_u64[h][slot<M>(idx)] = st[h][q];
            }
    } else if (xlast && near1) {
        // done by the x pass (XUPD)
    } else if (lsrk) {
#pragma unroll
        for (int h = 0; h < V; ++h) {
            double reg[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = tid + q * NT;
                reg[q] = (idx < nvalid_nodes && s != 1) ? a.r0[h * EM3 + e0 * M3 + idx] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int idx = tid + q * NT;
                if (idx < nvalid_nodes) {
                    const long long gi = h * EM3 + e0 * M3 + idx;
                    const int sl = slot<M>(idx);
                    s_u64[h][sl] = probe_lsrk(a, gi, s, dt, s_u64[h][sl], reg[q], dudt(h, sl, 0.0));
                }
            }
        }
    } else {
#pragma unroll
        for (int h = 0; h < V; ++h)
            for (int idx = tid; idx < nvalid_nodes; idx += NT) {
                const long long gi = h * EM3 + e0 * M3 + idx;
                const int sl = slot<M>(idx);
                s_u64[h][sl] = probe(gi, s_u64[h][sl], dudt(h, sl, 0.0));
            }
    }
    __syncthreads();

    // ---- positivity repair (kernel_repair after the probe) ---------------
    if (a.repair) {
        unsigned bad = 0;
        for (int idx = tid; idx < nvalid_nodes; idx += NT) {
            double uu[V];
#pragma unroll
            for (int h = 0; h < V; ++h) uu[h] = s_u64[h][slot<M>(idx)];
            if (!isvalid64<EQ>(a.ph64, uu)) bad |= 1u << (idx / M3);
        }
        unsigned bad_el = 0;
#pragma unroll
        for (int el = 0; el < BE; ++el)
            if (__any((bad >> el) & 1u)) bad_el |= 1u << el;
        if (bad_el) {
            if (tid < BE && ((bad_el >> tid) & 1u)) {
                double* su[V];
#pragma unroll
                for (int h = 0; h < V; ++h) su[h] = s_u64[h];
                if (!repair_element<EQ, M>(a.ph64, a.rk.w64, su, tid * M3)) atomicOr(a.status, 1);
            }
            __syncthreads();
        }
    }

    // ---- write the new state; SSP register r1 takes the repaired state ---
#pragma unroll
    for (int h = 0; h < V; ++h)
        for (int idx = tid; idx < nvalid_nodes; idx += NT) {
            const long long gi = h * EM3 + e0 * M3 + idx;
            const double st = s_u64[h][slot<M>(idx)];
            if constexpr (COMPACT)
                __builtin_nontemporal_store(st, &a.u_out[gi]);
            else
                a.u_out[gi] = st;
            if (a.timedisc == TD_SSPRK54) {
                if (s == 2) a.r1[gi] = a.rk.ssp[20] * st;
                else if (s == 3) a.r1[gi] = a.r1[gi] + a.rk.ssp[21] * st;
            }
        }

    // ---- CFL partials of the new state (last stage of a step) -----------
    if (a.cfl_part) block_cfl<EQ, M, BE, NT>(a.ph64, a.delta, e0, a.nelem, s_u64, &s_raw[0], a.cfl_part + 4 * blk);
}

// Minimum waves per SIMD the register allocation must leave room for.
// FP64 Euler at N=4: left to itself the compiler takes 256 VGPRs and a few
// AGPRs, one wave per SIMD, although the 20 KB of LDS allow two; asked for
// two it fits in at most 245 VGPRs without spills and runs 1.5x faster.
// At N=3 the request keeps the two paths of the log-means (with and without
// the log's range reduction) within 256 VGPRs as well.
// (The same request for three waves at N=2 spills and runs slower.)
template <int EQ, int M, typename Real>
constexpr int stage_waves_per_eu() {
    return (EQ == EQ_EULER && M >= 4 && sizeof(Real) == 8) ? 2 : 1;
}

// One workgroup per block of BE elements.
template <int EQ, int M, int BE, int VOL, int SURF, typename Real, int IK = IK_RUNTIME>
__global__ __launch_bounds__(BE * M * M) __attribute__((amdgpu_waves_per_eu(stage_waves_per_eu<EQ, M, Real>())))
void stage_kernel(StageArgs<M, Real> a) {
    __shared__ __align__(16) double lds[StageLds<EQ, M, BE, Real>::DOUBLES];
    const long long nblocks = (a.nelem + BE - 1) / BE;
    stage_block<EQ, M, BE, VOL, SURF, Real, IK>(a, xcd_block(nblocks), lds);
}

// Persistent form for the batched ensemble: a grid of gridDim.x workgroups
// walks all blocks; with xcd != 0 the walk is dealt so that each XCD (the
// hardware deals workgroups round-robin over the 8 XCDs) takes a contiguous
// stretch of blocks.
template <int EQ, int M, int BE, int VOL, int SURF, typename Real>
__global__ __launch_bounds__(BE * M * M) void stage_kernel_persistent(StageArgs<M, Real> a, int xcd) {
    __shared__ __align__(16) double lds[StageLds<EQ, M, BE, Real>::DOUBLES];
    const long long nblocks = (a.nelem + BE - 1) / BE;
    const long long G = gridDim.x;
    const long long per = (nblocks + G - 1) / G;
    for (long long it = 0; it < per; ++it) {
        long long b;
        if (xcd && (G & 7) == 0) {
            // workgroup w sits on XCD w % 8; XCD x owns blocks [x*nb8, (x+1)*nb8)
            const long long w = blockIdx.x, x = w & 7, lw = w >> 3, G8 = G >> 3;
            const long long nb8 = (nblocks + 7) / 8;
            b = x * nb8 + it * G8 + lw;
            if (b >= (x + 1) * nb8) continue;
        } else {
            b = it * G + blockIdx.x;
        }
        if (b >= nblocks) continue;
        stage_block<EQ, M, BE, VOL, SURF, Real, IK_RUNTIME, false>(a, b, lds);
        __syncthreads();
    }
}

// Standalone CFL reduction of a state in HBM (first step, calc_timestep):
// the same per-block partials as the fused stage epilogue.
template <int EQ, int M, int BE>
__global__ __launch_bounds__(BE * M * M) void cfl_kernel(const double* __restrict__ u, const double* __restrict__ delta,
                                                         long long nelem, Phys64 P, double* __restrict__ part) {
    constexpr int M3 = M * M * M, NN = BE * M3, NT = BE * M * M, V = NVars<EQ>::V;
    __shared__ double s_u[V][NN];
    __shared__ double s_w[4][NT];
    const int tid = threadIdx.x;
    const long long blk = blockIdx.x;
    const long long e0 = blk * BE;
    const long long EM3 = nelem * (long long)M3;
    const long long nvalid = (nelem - e0 < BE ? nelem - e0 : BE) * (long long)M3;
#pragma unroll
    for (int h = 0; h < V; ++h)
        for (int idx = tid; idx < NN; idx += NT)
            s_u[h][slot<M>(idx)] = idx < nvalid ? u[h * EM3 + e0 * M3 + idx] : (h == 0 || h == 4 ? 1.0 : 0.0);
    __syncthreads();
    block_cfl<EQ, M, BE, NT>(P, delta, e0, nelem, s_u, &s_w[0][0], part + 4 * blk);
}

// Deterministic reduction of the per-block partials (min and max are exact,
// so the result does not depend on the order) and
// dt = 0.5*CFL/(2*N_NODES-1) * dtmin (kernel_timestep_mod.f90:49); for
// GLM-MHD also c_h = 2|lmax - vmax| and alpha = c_h/0.18
// (kernel/mcfv/3d/kernel_timestep_mod.f90:53-56).
// dt_dev: [0] dt, [1] lmax, [2] vmax, [3] c_h, [4] alpha.
// Two levels: block i of a grid of G <= CFL_GRID blocks reduces the i-th
// contiguous G-th of the block partials (3 of the 4 values each: dt
// candidate, lmax, vmax) into red[i]; the block that finishes last (ticket) reduces red and
// writes dt_dev (and hist[slot]), then rearms the ticket for the next call.
// min and max are exact, so the result does not depend on the order.
constexpr int CFL_GRID = 1024, CFL_NT = 256;

__device__ __forceinline__ void cfl_block_reduce(double& dt, double& lm, double& vm, double (*s)[CFL_NT]) {
    const int tid = threadIdx.x;
    s[0][tid] = dt;
    s[1][tid] = lm;
    s[2][tid] = vm;
    __syncthreads();
    for (int off = CFL_NT / 2; off > 0; off >>= 1) {
        if (tid < off) {
            s[0][tid] = s[0][tid + off] < s[0][tid] ? s[0][tid + off] : s[0][tid];
            s[1][tid] = s[1][tid + off] > s[1][tid] ? s[1][tid + off] : s[1][tid];
            s[2][tid] = s[2][tid + off] > s[2][tid] ? s[2][tid + off] : s[2][tid];
        }
        __syncthreads();
    }
    dt = s[0][0];
    lm = s[1][0];
    vm = s[2][0];
}

static __global__ __launch_bounds__(CFL_NT) void cfl_finalize(const double* __restrict__ part, long long npart,
                                                              double cfl, int M, double* __restrict__ dt_dev,
                                                              double* __restrict__ hist, long long slot,
                                                              double* __restrict__ red, unsigned int* __restrict__ ticket) {
    __shared__ double s[3][CFL_NT];
    __shared__ bool last;
    const int tid = threadIdx.x;
    const long long G = gridDim.x;
    const long long chunk = (npart + G - 1) / G, i0 = blockIdx.x * chunk;
    const long long i1 = i0 + chunk < npart ? i0 + chunk : npart;
    double dt = 1.7976931348623157e308, lm = 0.0, vm = 0.0;
    for (long long i = i0 + tid; i < i1; i += CFL_NT) {
        const double* q = part + 4 * i;
        dt = q[0] < dt ? q[0] : dt;
        lm = q[1] > lm ? q[1] : lm;
        vm = q[2] > vm ? q[2] : vm;
    }
    cfl_block_reduce(dt, lm, vm, s);
    if (tid == 0) {
        red[3 * blockIdx.x + 0] = dt;
        red[3 * blockIdx.x + 1] = lm;
        red[3 * blockIdx.x + 2] = vm;
        __threadfence();
        last = atomicAdd(ticket, 1u) == unsigned(G - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    dt = 1.7976931348623157e308;
    lm = vm = 0.0;
    for (long long i = tid; i < G; i += CFL_NT) {
        const volatile double* q = red + 3 * i;
        dt = q[0] < dt ? q[0] : dt;
        lm = q[1] > lm ? q[1] : lm;
        vm = q[2] > vm ? q[2] : vm;
    }
    cfl_block_reduce(dt, lm, vm, s);
    if (tid == 0) {
        const double d = 0.5 * cfl / double(2 * M - 1) * dt;
        dt_dev[0] = d;
        dt_dev[1] = lm;
        dt_dev[2] = vm;
        const double ch = 2.0 * fabs(lm - vm);
        dt_dev[3] = ch;
        dt_dev[4] = ch / 0.18;
        if (hist && slot >= 0) hist[slot] = d;
        *ticket = 0u;
    }
}

}  // namespace nemo
