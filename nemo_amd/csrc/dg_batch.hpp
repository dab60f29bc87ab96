// Batched ensemble: one launch per RK stage for the blocks of every
// polynomial degree (gfx950).
//
// The runs of the ensemble are grouped by degree (M = 3, 4, 5 nodes per
// direction); each group keeps its own arrays and block tile (BE elements of
// M^2 lanes in one wavefront: 7 x 9, 4 x 16, 2 x 25).  A launch covers the
// blocks of all groups: its block range is the concatenation of the group
// ranges in a chosen order, and a workgroup takes its degree and group-local
// block from its index, uniformly (scalar), then runs the stage of that
// block with the group's arguments.  The workgroup is one wavefront of 64
// lanes; the tiles of M = 3 and M = 5 leave 1 and 14 of them idle.  The
// stage arguments of the three groups live in device memory (their kernel
// argument form would exceed the kernarg segment).
#pragma once

#include "dg_stage.hpp"

namespace nemo {

template <int M>
struct BatchTile;
template <>
struct BatchTile<3> { static constexpr int BE = 7; };
template <>
struct BatchTile<4> { static constexpr int BE = 4; };
template <>
struct BatchTile<5> { static constexpr int BE = 2; };

template <typename Real>
struct MultiStageArgs {
    StageArgs<3, Real> a3;
    StageArgs<4, Real> a4;
    StageArgs<5, Real> a5;
    long long nblocks[3];  // blocks of degree M = 3, 4, 5
    int order[3];          // launch order of the groups (group index M - 3)
};

template <typename Real>
struct MultiLds {
    static constexpr int D3 = StageLds<EQ_EULER, 3, BatchTile<3>::BE, Real>::DOUBLES;
    static constexpr int D4 = StageLds<EQ_EULER, 4, BatchTile<4>::BE, Real>::DOUBLES;
    static constexpr int D5 = StageLds<EQ_EULER, 5, BatchTile<5>::BE, Real>::DOUBLES;
    static constexpr int DOUBLES = D3 > D4 ? (D3 > D5 ? D3 : D5) : (D4 > D5 ? D4 : D5);
};

// Group-local block b of group g (0, 1, 2 = M 3, 4, 5) through a whole stage.
template <typename Real>
__device__ __forceinline__ void multi_block(const MultiStageArgs<Real>& ma, int g, long long b, double* lds) {
    const int tid = threadIdx.x;
    if (g == 0) {
        if (tid < BatchTile<3>::BE * 9)
            stage_block<EQ_EULER, 3, BatchTile<3>::BE, VOL_EC, SURF_HLLC, Real, IK_RUNTIME, false>(ma.a3, b, lds);
    } else if (g == 1) {
        stage_block<EQ_EULER, 4, BatchTile<4>::BE, VOL_EC, SURF_HLLC, Real, IK_RUNTIME, false>(ma.a4, b, lds);
    } else {
        if (tid < BatchTile<5>::BE * 25)
            stage_block<EQ_EULER, 5, BatchTile<5>::BE, VOL_EC, SURF_HLLC, Real, IK_RUNTIME, false>(ma.a5, b, lds);
    }
}

// Waves per SIMD the register allocation must leave room for: the FP64
// stage code of M = 4 and 5 fits two waves (as the single-run kernels do);
// left alone the compiler gives the combined kernel 256 VGPRs and one wave.
template <typename Real>
constexpr int multi_waves_per_eu() {
    return sizeof(Real) == 8 ? 2 : 1;
}

// One workgroup per block: workgroup index -> (group, block).  Within each
// group's stretch of indices the blocks are dealt to the 8 XCDs in
// contiguous stretches (xcd_block), as in the single-run launch.
template <typename Real>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(multi_waves_per_eu<Real>())))
void stage_kernel_multi(const MultiStageArgs<Real>* __restrict__ pma) {
    __shared__ __align__(16) double lds[MultiLds<Real>::DOUBLES];
    const MultiStageArgs<Real>& ma = *pma;
    long long w = blockIdx.x;
    int g = ma.order[0];
    for (int k = 0; k < 3; ++k) {
        g = ma.order[k];
        if (w < ma.nblocks[g]) break;
        w -= ma.nblocks[g];
    }
    const long long nb = ma.nblocks[g];
    long long b = w;
    if ((nb & 7) == 0) b = (w & 7) * (nb >> 3) + (w >> 3);
    multi_block<Real>(ma, g, b, lds);
}

// Persistent form: a grid of gridDim.x workgroups (a multiple of 8 with the
// XCD map) walks the groups in launch order; with xcd != 0 workgroup w (on
// XCD w % 8) takes the blocks of its XCD's contiguous eighth of each group.
// All loop bounds are uniform, so every wave leaves every loop.
template <typename Real>
__global__ __launch_bounds__(64) void stage_kernel_multi_persistent(const MultiStageArgs<Real>* __restrict__ pma,
                                                                    int xcd) {
    __shared__ __align__(16) double lds[MultiLds<Real>::DOUBLES];
    const MultiStageArgs<Real>& ma = *pma;
    const long long G = gridDim.x, w = blockIdx.x;
    for (int k = 0; k < 3; ++k) {
        const int g = ma.order[k];
        const long long nb = ma.nblocks[g];
        if (xcd && (G & 7) == 0) {
            const long long x = w & 7, lw = w >> 3, G8 = G >> 3;
            const long long nb8 = (nb + 7) / 8;
            for (long long j = lw; j < nb8; j += G8) {
                const long long b = x * nb8 + j;
                if (b >= nb) break;
                multi_block<Real>(ma, g, b, lds);
                __syncthreads();
            }
        } else {
            for (long long b = w; b < nb; b += G) {
                multi_block<Real>(ma, g, b, lds);
                __syncthreads();
            }
        }
    }
}

}  // namespace nemo
