// Non-conforming (2:1) faces, boundary faces and the state transfer of mesh
// adaptation (gfx950).
//
// mortar_kernel: the reference's mortar sequence of
// kernel_fill_faceflux_cb (kernel/dgsem/weakform/2d/kernel_flux_mod.f90:
// 58-147, 3-D face arrays as in kernel/mcfv/3d/kernel_flux_mod.f90:100-121):
// the big element's face values are interpolated to the nodes of each of the
// four half-size faces (rfaceMat), the Riemann problem is solved at those
// nodes, each small element takes its fluxes as they are and the big element
// takes their projection (cfaceMat) summed over the four.  GLM-MHD adds the
// Powell/GLM term of each side before the projection (kernel/mcfv/3d/
// kernel_flux_mod.f90:229-268): the side's inner state at the half-face node
// (the small element's own node, the big element's interpolated one) times
// the face averages of B_n and psi.  One workgroup per mortar, one lane per
// small-face node; the face fluxes go to slots of a small buffer that the
// stage kernel reads in place of its own face flux on those faces.
//
// remap_kernel: kernel_amr_refine / kernel_amr_coarsen
// (kernel/mcfv/3d/kernel_amr_mod.f90:12-70) with the positivity fallback of
// setup_amr_replace_refine/coarsen (setups/euler/3d/sedov/setup_amr_mod.f90:
// 76-130; for GLM-MHD with that equation set's density and gas pressure):
// one workgroup per replacement unit (copy, one parent into eight children,
// eight children into one parent), one lane per node.
#pragma once

#include "dg_euler.hpp"

namespace nemo {

template <int M>
struct MortarOps {
    double P[2][M][M];  // P[c][i][p]: parent node p -> node i of half c
    double R[2][M][M];  // R[c][p][i]: node i of half c -> parent node p (projection)
};

template <int M>
struct MortarShape {
    static constexpr int M2 = M * M;
    static constexpr int NT = (4 * M2 + 63) / 64 * 64;
};

template <int M, typename Real>
struct MortarArgs {
    const double* __restrict__ u_in;     // [var][elem][node]
    double* __restrict__ flux;           // [slot][var][M*M]
    const int32_t* __restrict__ mortars; // [nmortar][6]: big, face, small_0..3
    const double* __restrict__ glm;      // GLM-MHD: c_h of the current step in glm[0]
    long long nelem;
    Gas<Real> gas;
    MhdGas<Real> mhd;
    Phys64 ph64;
    MortarOps<M> op;
};

// Node index of face node (ta, tb) at line position l of direction d.
template <int M>
__device__ __forceinline__ int face_node(int d, int l, int ta, int tb) {
    return d == 0 ? nidx<0, M>(l, ta, tb) : (d == 1 ? nidx<1, M>(l, ta, tb) : nidx<2, M>(l, ta, tb));
}

// Face flux of one half-face node for each side: fs for the small element
// (uL or uR per big_left), fb for the big one; equal for Euler.
template <int EQ, int SURF, int D, typename Real>
__device__ __forceinline__ void mortar_riemann_d(const Gas<Real>& g, const MhdGas<Real>& mg, const Phys64& P,
                                                 const double* uL, const double* uR, bool big_left, Real* fs,
                                                 Real* fb) {
    if constexpr (EQ == EQ_MHD) {
        constexpr int BN = Mag<D, 0>::idx;
        MhdState<Real> L, R;
#pragma unroll
        for (int h = 0; h < 9; ++h) {
            L.u[h] = Real(uL[h]);
            R.u[h] = Real(uR[h]);
        }
        mhd_make_state<false>(mg, L);
        mhd_make_state<false>(mg, R);
        if constexpr (SURF == SURF_HLLD)
            mhd_hlld<D>(mg, L, R, fs);
        else
            mhd_rusanov<D>(mg, L, R, fs);
#pragma unroll
        for (int h = 0; h < 9; ++h) fb[h] = fs[h];
        const Real Bavg = Real(0.5) * (L.u[BN] + R.u[BN]);
        const Real Gavg = Real(0.5) * (L.u[GLMP] + R.u[GLMP]);
        mhd_nc_add<D>(mg, big_left ? R : L, Bavg, Gavg, fs);
        mhd_nc_add<D>(mg, big_left ? L : R, Bavg, Gavg, fb);
    } else {
        NodeState<Real> L, R;
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            L.u[h] = Real(uL[h]);
            R.u[h] = Real(uR[h]);
        }
        if constexpr (sizeof(Real) == 4) {
            L.p64 = pressure(P.gas, uL);
            R.p64 = pressure(P.gas, uR);
        }
        make_state<D, false>(g, L);
        make_state<D, false>(g, R);
        riemann<SURF, D>(g, L, R, fs);
#pragma unroll
        for (int h = 0; h < 5; ++h) fb[h] = fs[h];
    }
}

template <int EQ, int M, int SURF, typename Real>
__global__ void __launch_bounds__(MortarShape<M>::NT) mortar_kernel(MortarArgs<M, Real> a) {
    constexpr int M2 = M * M, M3 = M2 * M, V = NVars<EQ>::V;
    __shared__ double s_big[V][M2];
    __shared__ double s_f[V][4 * M2];
    const int tid = threadIdx.x;
    const long long m = blockIdx.x;
    const int32_t* mt = a.mortars + m * 6;
    const long long big = mt[0];
    const int fb = mt[1];
    const int d = fb >> 1;
    const bool big_left = (fb & 1) != 0;  // big face is x+/y+/z+: big state is u_L
    const long long EM3 = a.nelem * (long long)M3;
    MhdGas<Real> mg = a.mhd;
    if constexpr (EQ == EQ_MHD) mg.ch = Real(a.glm[0]);
    if (tid < M2) {
        const int n = face_node<M>(d, big_left ? M - 1 : 0, tid % M, tid / M);
#pragma unroll
        for (int h = 0; h < V; ++h) s_big[h][tid] = a.u_in[h * EM3 + big * M3 + n];
    }
    __syncthreads();
    if (tid < 4 * M2) {
        const int k = tid / M2, t = tid - k * M2;
        const int ta = t % M, tb = t / M, ka = k & 1, kb = k >> 1;
        const long long sm = mt[2 + k];
        const int n = face_node<M>(d, big_left ? 0 : M - 1, ta, tb);
        double us[V], ub[V];
#pragma unroll
        for (int h = 0; h < V; ++h) {
            us[h] = a.u_in[h * EM3 + sm * M3 + n];
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < M; ++q) {
                double r = 0.0;
#pragma unroll
                for (int p = 0; p < M; ++p) r += a.op.P[ka][ta][p] * s_big[h][p + M * q];
                acc += a.op.P[kb][tb][q] * r;
            }
            ub[h] = acc;
        }
        const double* uL = big_left ? ub : us;
        const double* uR = big_left ? us : ub;
        Real fsm[V], fbg[V];
        if (d == 0)
            mortar_riemann_d<EQ, SURF, 0>(a.gas, mg, a.ph64, uL, uR, big_left, fsm, fbg);
        else if (d == 1)
            mortar_riemann_d<EQ, SURF, 1>(a.gas, mg, a.ph64, uL, uR, big_left, fsm, fbg);
        else
            mortar_riemann_d<EQ, SURF, 2>(a.gas, mg, a.ph64, uL, uR, big_left, fsm, fbg);
        double* out = a.flux + (size_t(5 * m + 1 + k) * V) * M2 + t;
#pragma unroll
        for (int h = 0; h < V; ++h) {
            out[h * M2] = double(fsm[h]);
            s_f[h][tid] = double(fbg[h]);
        }
    }
    __syncthreads();
    if (tid < M2) {
        const int p = tid % M, q = tid / M;
        double* out = a.flux + (size_t(5 * m) * V) * M2 + tid;
#pragma unroll
        for (int h = 0; h < V; ++h) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int ka = k & 1, kb = k >> 1;
#pragma unroll
                for (int tb = 0; tb < M; ++tb) {
                    double r = 0.0;
#pragma unroll
                    for (int ta = 0; ta < M; ++ta) r += a.op.R[ka][p][ta] * s_f[h][k * M2 + ta + M * tb];
                    acc += a.op.R[kb][q][tb] * r;
                }
            }
            out[h * M2] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Boundary faces of non-periodic meshes
// ---------------------------------------------------------------------------

enum { BC_TRANSMISSIVE = 0, BC_SLIP_WALL = 1 };

template <int M>
struct BoundaryArgs {
    const double* __restrict__ u_in;     // [var][elem][node]
    double* __restrict__ flux;           // [slot][var][M*M]
    const int32_t* __restrict__ bfaces;  // [nbound][3]: element, face, kind
    const double* __restrict__ glm;      // GLM-MHD: c_h of the current step in glm[0]
    long long nelem, nbound, slot0;      // slot0 = 5 nmortar: the boundary slots follow the mortars'
    Phys64 ph64;
};

// Wall pressure of the Eulerian slip wall (equations/euler/riemann/
// eulerian-slip-wall/riemann_mod.f90:8-38): the Riemann problem between the
// inner state and its mirror image, a rarefaction for v <= 0 (v: velocity
// toward the wall; the base floored at 1e-5 as there) and a shock otherwise.
__device__ __forceinline__ double slip_wall_pressure(double kappa, double rho, double v, double p) {
    if (v <= 0.0) {
        const double c = sqrt(kappa * p / rho);
        return p * pow(fmax(1.0 + 0.5 * (kappa - 1.0) * v / c, 1e-5), 2.0 * kappa / (kappa - 1.0));
    }
    const double a = 2.0 / (kappa + 1.0) / rho, b = (kappa - 1.0) / (kappa + 1.0) * p;
    return p + v / a * 0.5 * (v + sqrt(v * v + 4.0 * a * (p + b)));
}

// Face flux of one boundary face node in direction D, FP64: the configured
// Riemann solver between the inner state and its copy (transmissive; for
// GLM-MHD with the inner state's non-conservative term), or the wall
// pressure in the normal momentum (slip wall, Euler).
template <int EQ, int SURF, int D>
__device__ __forceinline__ void boundary_flux_d(const Phys64& P, double ch, const double* ui, int kind, bool plus,
                                                double* f) {
    if constexpr (EQ == EQ_MHD) {
        constexpr int BN = Mag<D, 0>::idx;
        MhdGas<double> mg = P.mhd;
        mg.ch = ch;
        MhdState<double> S;
#pragma unroll
        for (int h = 0; h < 9; ++h) S.u[h] = ui[h];
        mhd_make_state<false>(mg, S);
        if constexpr (SURF == SURF_HLLD)
            mhd_hlld<D>(mg, S, S, f);
        else
            mhd_rusanov<D>(mg, S, S, f);
        mhd_nc_add<D>(mg, S, S.u[BN], S.u[GLMP], f);
    } else {
        if (kind == BC_SLIP_WALL) {
            const double v = (plus ? 1.0 : -1.0) * ui[1 + D] / ui[DENS];
#pragma unroll
            for (int h = 0; h < 5; ++h) f[h] = 0.0;
            f[1 + D] = slip_wall_pressure(P.gas.kappa, ui[DENS], v, pressure(P.gas, ui));
        } else {
            NodeState<double> S;
#pragma unroll
            for (int h = 0; h < 5; ++h) S.u[h] = ui[h];
            make_state<D, false>(P.gas, S);
            riemann<SURF, D>(P.gas, S, S, f);
        }
    }
}

// One lane per boundary face node: the inner node's state, the face flux,
// the face's slot (read by the stage kernel as for a mortar face).
template <int EQ, int M, int SURF>
__global__ void __launch_bounds__(64) boundary_kernel(BoundaryArgs<M> a) {
    constexpr int M2 = M * M, M3 = M2 * M, V = NVars<EQ>::V;
    const long long t = (long long)blockIdx.x * 64 + threadIdx.x;
    if (t >= a.nbound * M2) return;
    const long long b = t / M2;
    const int q = int(t - b * M2), ta = q % M, tb = q / M;
    const int32_t* bf = a.bfaces + b * 3;
    const long long e = bf[0];
    const int face = bf[1], kind = bf[2], d = face >> 1;
    const bool plus = (face & 1) != 0;
    const long long EM3 = a.nelem * (long long)M3;
    const int n = face_node<M>(d, plus ? M - 1 : 0, ta, tb);
    double ui[V], f[V];
#pragma unroll
    for (int h = 0; h < V; ++h) ui[h] = a.u_in[h * EM3 + e * M3 + n];
    const double ch = EQ == EQ_MHD ? a.glm[0] : 0.0;
    if (d == 0)
        boundary_flux_d<EQ, SURF, 0>(a.ph64, ch, ui, kind, plus, f);
    else if (d == 1)
        boundary_flux_d<EQ, SURF, 1>(a.ph64, ch, ui, kind, plus, f);
    else
        boundary_flux_d<EQ, SURF, 2>(a.ph64, ch, ui, kind, plus, f);
    double* out = a.flux + size_t(a.slot0 + b) * V * M2 + q;
#pragma unroll
    for (int h = 0; h < V; ++h) out[h * M2] = f[h];
}

// ---------------------------------------------------------------------------
// State transfer of one adaptation (FP64, Euler or GLM-MHD)
// ---------------------------------------------------------------------------

enum { UNIT_COPY = 0, UNIT_REFINE = 1, UNIT_COARSEN = 2 };

template <int M>
struct RemapShape {
    static constexpr int M3 = M * M * M;
    static constexpr int NT = (M3 + 63) / 64 * 64;
};

template <int M>
struct RemapArgs {
    const double* __restrict__ u_old;    // [var][old elem][node]
    double* __restrict__ u_new;          // [var][new elem][node]
    const int32_t* __restrict__ units;   // [nunits][3]: kind, first old, first new
    long long nold, nnew;
    Phys64 ph64;
    MortarOps<M> op;
    double w[M];
    int* __restrict__ fallbacks;         // count of units that fell back to the mean
};

// Block sum of v over the NT lanes (result valid in every lane).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* s_red) {
    const int tid = threadIdx.x;
    s_red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int off = NT / 2; off > 0; off >>= 1) {
        if (tid < off) s_red[tid] += s_red[tid + off];
        __syncthreads();
    }
    const double r = s_red[0];
    __syncthreads();
    return r;
}

// The variables pass through LDS in batches of VB (all five for Euler, as
// one batch: every load of a unit is issued before the first use; three for
// GLM-MHD, which keeps the eight children of a coarsening unit within 24 KB
// at N = 4); the results of all variables stay in registers until the
// unit's validity is known.
template <int EQ>
struct RemapBatch {
    static constexpr int VB = EQ == EQ_MHD ? 3 : 5;
};

template <int EQ, int M>
__global__ void __launch_bounds__(RemapShape<M>::NT) remap_kernel(RemapArgs<M> a) {
    constexpr int M3 = M * M * M, NT = RemapShape<M>::NT, V = NVars<EQ>::V, VB = RemapBatch<EQ>::VB;
    static_assert(V % VB == 0, "whole batches");
    __shared__ double s_in[8][VB][M3];
    __shared__ double s_red[NT];
    const int tid = threadIdx.x;
    const int32_t* ut = a.units + blockIdx.x * 3;
    const int kind = ut[0];
    const long long io = ut[1], jn = ut[2];
    const long long OM3 = a.nold * (long long)M3, NM3 = a.nnew * (long long)M3;
    const bool act = tid < M3;
    const int i = tid % M, j = (tid / M) % M, k = tid / (M * M);
    const double wn = act ? 0.125 * a.w[i] * a.w[j] * a.w[k] : 0.0;

    if (kind == UNIT_COPY) {
        if (act)
#pragma unroll
            for (int h = 0; h < V; ++h) a.u_new[h * NM3 + jn * M3 + tid] = a.u_old[h * OM3 + io * M3 + tid];
        return;
    }

    if (kind == UNIT_REFINE) {
        // child c = cx + 2 cy + 4 cz at new element jn + c
        double out[8][V];
#pragma unroll
        for (int h0 = 0; h0 < V; h0 += VB) {
            __syncthreads();  // the previous batch's reads are done
            if (act)
#pragma unroll
                for (int hb = 0; hb < VB; ++hb) s_in[0][hb][tid] = a.u_old[(h0 + hb) * OM3 + io * M3 + tid];
            __syncthreads();
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
#pragma unroll
                for (int hb = 0; hb < VB; ++hb) {
                    double acc = 0.0;
                    if (act)
                        for (int kk = 0; kk < M; ++kk) {
                            double sj = 0.0;
                            for (int jj = 0; jj < M; ++jj) {
                                double si = 0.0;
#pragma unroll
                                for (int ii = 0; ii < M; ++ii)
                                    si += a.op.P[cx][i][ii] * s_in[0][hb][ii + M * jj + M * M * kk];
                                sj += a.op.P[cy][j][jj] * si;
                            }
                            acc += a.op.P[cz][k][kk] * sj;
                        }
                    out[c][h0 + hb] = acc;
                }
            }
        }
        bool ok = true;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (act) ok = ok && isvalid64<EQ>(a.ph64, out[c]);
        const bool bad = __syncthreads_or(!ok);
        double mean[V];
        if (bad) {
#pragma unroll
            for (int h = 0; h < V; ++h)
                mean[h] = block_sum<NT>(act ? wn * a.u_old[h * OM3 + io * M3 + tid] : 0.0, s_red);
            if (tid == 0) atomicAdd(a.fallbacks, 1);
        }
        if (act)
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int h = 0; h < V; ++h) a.u_new[h * NM3 + (jn + c) * M3 + tid] = bad ? mean[h] : out[c][h];
        return;
    }

    // UNIT_COARSEN: children io .. io+7 (child order) into new element jn
    double out[V];
#pragma unroll
    for (int h0 = 0; h0 < V; h0 += VB) {
        __syncthreads();
        if (act)
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int hb = 0; hb < VB; ++hb) s_in[c][hb][tid] = a.u_old[(h0 + hb) * OM3 + (io + c) * M3 + tid];
        __syncthreads();
#pragma unroll
        for (int hb = 0; hb < VB; ++hb) {
            double acc = 0.0;
            if (act)
                for (int c = 0; c < 8; ++c) {
                    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                    for (int kk = 0; kk < M; ++kk) {
                        double sj = 0.0;
                        for (int jj = 0; jj < M; ++jj) {
                            double si = 0.0;
#pragma unroll
                            for (int ii = 0; ii < M; ++ii)
                                si += a.op.R[cx][i][ii] * s_in[c][hb][ii + M * jj + M * M * kk];
                            sj += a.op.R[cy][j][jj] * si;
                        }
                        acc += a.op.R[cz][k][kk] * sj;
                    }
                }
            out[h0 + hb] = acc;
        }
    }
    const bool ok = !act || isvalid64<EQ>(a.ph64, out);
    const bool bad = __syncthreads_or(!ok);
    if (bad) {
#pragma unroll
        for (int h = 0; h < V; ++h) out[h] = block_sum<NT>(act ? wn * out[h] : 0.0, s_red);
        if (tid == 0) atomicAdd(a.fallbacks, 1);
    }
    if (act)
#pragma unroll
        for (int h = 0; h < V; ++h) a.u_new[h * NM3 + jn * M3 + tid] = out[h];
}

}  // namespace nemo
