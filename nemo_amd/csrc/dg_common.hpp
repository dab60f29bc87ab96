// Shared definitions of the gfx950 stage kernels: operator tables, kernel
// arguments, node indexing and the LDS layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "physics.hpp"
#include "physics_mhd.hpp"

namespace nemo {

enum { EQ_EULER = 0, EQ_MHD = 1 };
enum { VOL_WEAK = 0, VOL_CENTRAL = 1, VOL_EC = 2 };
enum { SURF_RUSANOV = 0, SURF_HLLC = 1, SURF_HLLD = 2 };
enum { TD_LSRK54 = 0, TD_SSPRK54 = 1 };
enum { IND_NONE = 0, IND_PP = 1, IND_MLP = 2 };
// How a stage kernel obtains the blending factor (compile time):
// IK_RUNTIME: none or Persson-Peraire by StageArgs::indicator;
// IK_MLP: the learned indicator with its judge (first stage of a step,
// diagnostics, training data); IK_STORED: the factor kept in alpha_step
// by the step's first stage (stages 2..5 of the learned indicator), so
// that those stages carry neither the network nor the Persson-Peraire code.
enum { IK_RUNTIME = 0, IK_MLP = 1, IK_STORED = 2 };

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

// Number of conserved variables.
template <int EQ>
struct NVars {
    static constexpr int V = EQ == EQ_MHD ? 9 : 5;
};

// FP64 constants of both equation systems (CFL, validity, the FP64 parts of
// the FP32 operator).
struct Phys64 {
    Gas<double> gas;
    MhdGas<double> mhd;
};

template <int EQ>
__device__ __forceinline__ double pressure64(const Phys64& P, const double* u) {
    if constexpr (EQ == EQ_MHD)
        return mhd_pressure(P.mhd, u);
    else
        return pressure(P.gas, u);
}

template <int EQ>
__device__ __forceinline__ bool isvalid64(const Phys64& P, const double* u) {
    return (u[DENS] > 0.0) && (pressure64<EQ>(P, u) > 0.0);
}

template <int EQ, int D>
__device__ __forceinline__ double lambda_max64(const Phys64& P, const double* u) {
    if constexpr (EQ == EQ_MHD)
        return mhd_lambda_max<D>(P.mhd, u);
    else
        return lambda_max<D>(P.gas, u);
}

template <int M, typename Real>
struct StageArgs {
    const double* __restrict__ u_in;
    double* __restrict__ u_out;
    double* __restrict__ r0;
    double* __restrict__ r1;
    double* __restrict__ rhs_out;    // non-null: write the RHS, no update
    double* __restrict__ alpha_out;  // optional per-element blending factor
    const int32_t* __restrict__ nbr; // [nelem][6]; -(s+1): face flux from mortar_flux slot s
    const double* __restrict__ mortar_flux;  // [slot][var][M*M] fluxes of the non-conforming faces
    const double* __restrict__ delta;// [nelem][3]
    const double* __restrict__ dt_dev;
    const double* __restrict__ glm;  // MHD: [c_h, alpha] of the current step
    const double* __restrict__ dt_run;     // batched ensemble: time step per run
    const int32_t* __restrict__ block_run; // batched ensemble: run of each block
    int* __restrict__ status;
    double* __restrict__ cfl_part;   // non-null: per-block CFL partials of the new state
    // learned indicator
    const float* __restrict__ mlp;   // network parameters, inference layout
    double* __restrict__ alpha_step; // per element factor of the current step
    float* __restrict__ feat_out;    // non-null: write features [nelem][4] (training data)
    unsigned long long* __restrict__ mlp_stats;  // judged, overridden, fallback
    int alpha_refresh;               // 1: compute the factor (first stage of a step)
    int alpha_only;                  // 1 (Euler): write the factor to alpha_out and stop there
    uint32_t step_index;             // sampling key of the judge
    uint32_t sample_thresh;          // judged fraction * 2^32
    double judge_tol;                // PP takes over where it exceeds the prediction by this
    long long nelem;
    double dt;
    int stage;      // 1..5
    int timedisc;
    int indicator;
    int repair;
    double alpha_max, alpha_min, ind_threshold, ind_s;
    Gas<Real> gas;
    MhdGas<Real> mhd;   // c_h is filled in by the kernel from glm[0]
    Phys64 ph64;        // FP64 constants (CFL, validity, FP64 parts of the FP32 operator)
    OpTable<M, Real> op;
    RKCoeffs rk;
};

template <int D, int M>
__device__ __forceinline__ constexpr int nidx(int l, int a, int b) {
    return D == 0 ? l + M * a + M * M * b : (D == 1 ? a + M * l + M * M * b : a + M * b + M * M * l);
}

// LDS slot of node n = i + M*j + M*M*k inside an element.  For M = 4 the
// slot is 16k + 4(i^j) + (i^k): a bijection under which the 16 lines of one
// element in any direction pass (one node per line) fall into 16 distinct
// slots modulo 16, so one element's lanes never share a bank.  Without it
// the x- and y-passes hit only 4 bank pairs (16-way conflicts).  Other M
// keep the natural order.
template <int M>
__device__ __forceinline__ constexpr int swz(int n) {
    if constexpr (M == 4) {
        const int i = n & 3, j = (n >> 2) & 3, k = n >> 4;
        return 16 * k + 4 * (i ^ j) + (i ^ k);
    } else {
        return n;
    }
}

// LDS slot of node n of the block element whose slots start at base =
// el*M^3.  (Spreading the four elements of a block over different bank
// groups as well, by XORing 16*(el & 3) into the slot, removes the
// remaining bank conflicts but measured slower: the extra address
// arithmetic costs more than the conflicts.)
template <int M>
__device__ __forceinline__ constexpr int bslot(int base, int n) {
    return base + swz<M>(n);
}

// LDS slot of block-local node index idx = element*M^3 + n.
template <int M>
__device__ __forceinline__ constexpr int slot(int idx) {
    constexpr int M3 = M * M * M;
    return bslot<M>((idx / M3) * M3, idx % M3);
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

// A compile-time flag passed as an argument (selects a variant of a pass).
template <bool B>
struct BoolConst {
    static constexpr bool value = B;
};

template <int M, int BE>
struct StageShape {
    static constexpr int M2 = M * M, M3 = M2 * M;
    static constexpr int NT = BE * M2;  // threads per workgroup
    static constexpr int NN = BE * M3;  // nodes per workgroup
};

}  // namespace nemo
