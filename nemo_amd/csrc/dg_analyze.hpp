// Volume integrals of the state on the device (gfx950): the sums of
// hook_analyze (setups/euler/2d/_common_/setup_analyze_mod.f90:34-164,
// 3-D form of setups/euler/3d/convergence-test/setup_analyze_mod.f90:40-151)
// with the LGL quadrature volumes w_i w_j w_k dx dy dz / 8 of every node,
// plus the extrema a positivity monitor needs.  Deterministic: one partial
// per workgroup (fixed tree), reduced in a fixed order by one workgroup.
#pragma once

#include "dg_common.hpp"

namespace nemo {

// Layout of the result (NEMO_ANALYZE_NSUMS doubles).
enum {
    AN_CONS = 0,      // [0, V): integrals of the conserved variables
    AN_EKIN = 9,      // kinetic energy  |m|^2 / (2 rho)
    AN_EINT = 10,     // internal energy p / (kappa - 1)
    AN_EMAG = 11,     // magnetic energy |B|^2 / (2 mu0) (GLM-MHD)
    AN_VOLUME = 12,   // total volume
    AN_RHO_MIN = 13,  // minimum nodal density
    AN_P_MIN = 14,    // minimum nodal pressure
    AN_V_MAX = 15,    // maximum nodal speed |v|
    AN_NSUMS = 16,
    AN_NT = 256
};

template <int M>
struct AnalyzeArgs {
    const double* __restrict__ u;       // [var][elem][node]
    const double* __restrict__ delta;   // [elem][3]
    double* __restrict__ part;          // [nblocks][AN_NSUMS]
    long long nelem;
    Phys64 ph;
    double w[M];
};

template <int EQ, int M>
__global__ void __launch_bounds__(AN_NT) analyze_kernel(AnalyzeArgs<M> a) {
    constexpr int V = NVars<EQ>::V, M3 = M * M * M;
    __shared__ double s[AN_NSUMS][AN_NT];
    const int tid = threadIdx.x;
    const long long idx = (long long)blockIdx.x * AN_NT + tid;
    const long long EM3 = a.nelem * (long long)M3;
    double q[AN_NSUMS];
#pragma unroll
    for (int k = 0; k < AN_NSUMS; ++k) q[k] = 0.0;
    q[AN_RHO_MIN] = q[AN_P_MIN] = 1.7976931348623157e308;
    if (idx < EM3) {
        const long long e = idx / M3;
        const int n = int(idx - e * M3);
        const int i = n % M, j = (n / M) % M, k = n / (M * M);
        const double vol = 0.125 * a.w[i] * a.w[j] * a.w[k] * a.delta[e * 3] * a.delta[e * 3 + 1] * a.delta[e * 3 + 2];
        double uu[V];
#pragma unroll
        for (int h = 0; h < V; ++h) uu[h] = a.u[h * EM3 + idx];
#pragma unroll
        for (int h = 0; h < V; ++h) q[AN_CONS + h] = vol * uu[h];
        const double m2 = uu[MOMX] * uu[MOMX] + uu[MOMY] * uu[MOMY] + uu[MOMZ] * uu[MOMZ];
        const double p = pressure64<EQ>(a.ph, uu);
        q[AN_EKIN] = vol * 0.5 * m2 / uu[DENS];
        if constexpr (EQ == EQ_MHD) {
            q[AN_EINT] = vol * p / a.ph.mhd.kappa_m1;
            q[AN_EMAG] = vol * 0.5 * (uu[MAGX] * uu[MAGX] + uu[MAGY] * uu[MAGY] + uu[MAGZ] * uu[MAGZ]) / a.ph.mhd.mu0;
        } else {
            q[AN_EINT] = vol * p / a.ph.gas.kappa_m1;
        }
        q[AN_VOLUME] = vol;
        q[AN_RHO_MIN] = uu[DENS];
        q[AN_P_MIN] = p;
        q[AN_V_MAX] = sqrt(m2) / uu[DENS];
    }
#pragma unroll
    for (int k = 0; k < AN_NSUMS; ++k) s[k][tid] = q[k];
    __syncthreads();
    for (int off = AN_NT / 2; off > 0; off >>= 1) {
        if (tid < off) {
#pragma unroll
            for (int k = 0; k < AN_RHO_MIN; ++k) s[k][tid] += s[k][tid + off];
            s[AN_RHO_MIN][tid] = fmin(s[AN_RHO_MIN][tid], s[AN_RHO_MIN][tid + off]);
            s[AN_P_MIN][tid] = fmin(s[AN_P_MIN][tid], s[AN_P_MIN][tid + off]);
            s[AN_V_MAX][tid] = fmax(s[AN_V_MAX][tid], s[AN_V_MAX][tid + off]);
        }
        __syncthreads();
    }
    if (tid < AN_NSUMS) a.part[(long long)blockIdx.x * AN_NSUMS + tid] = s[tid][0];
}

// Fixed-order reduction of the workgroup partials (one workgroup).
static __global__ void __launch_bounds__(AN_NT) analyze_finalize(const double* __restrict__ part, long long nblocks,
                                                                 double* __restrict__ out) {
    __shared__ double s[AN_NSUMS][AN_NT];
    const int tid = threadIdx.x;
    double q[AN_NSUMS];
#pragma unroll
    for (int k = 0; k < AN_NSUMS; ++k) q[k] = 0.0;
    q[AN_RHO_MIN] = q[AN_P_MIN] = 1.7976931348623157e308;
    for (long long b = tid; b < nblocks; b += AN_NT) {
        const double* p = part + b * AN_NSUMS;
#pragma unroll
        for (int k = 0; k < AN_RHO_MIN; ++k) q[k] += p[k];
        q[AN_RHO_MIN] = fmin(q[AN_RHO_MIN], p[AN_RHO_MIN]);
        q[AN_P_MIN] = fmin(q[AN_P_MIN], p[AN_P_MIN]);
        q[AN_V_MAX] = fmax(q[AN_V_MAX], p[AN_V_MAX]);
    }
#pragma unroll
    for (int k = 0; k < AN_NSUMS; ++k) s[k][tid] = q[k];
    __syncthreads();
    for (int off = AN_NT / 2; off > 0; off >>= 1) {
        if (tid < off) {
#pragma unroll
            for (int k = 0; k < AN_RHO_MIN; ++k) s[k][tid] += s[k][tid + off];
            s[AN_RHO_MIN][tid] = fmin(s[AN_RHO_MIN][tid], s[AN_RHO_MIN][tid + off]);
            s[AN_P_MIN][tid] = fmin(s[AN_P_MIN][tid], s[AN_P_MIN][tid + off]);
            s[AN_V_MAX][tid] = fmax(s[AN_V_MAX][tid], s[AN_V_MAX][tid + off]);
        }
        __syncthreads();
    }
    if (tid < AN_NSUMS) out[tid] = s[tid][0];
}

}  // namespace nemo
