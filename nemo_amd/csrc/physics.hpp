// Device physics for the compressible Euler equations (gfx950).
//
// Each function restates one reference function (paths relative to the
// reference's source/ directory).  Directions are compile-time: the
// reference's y/z variants rotate the momentum components into the x slot,
// call the x function and rotate back (equations/euler/equations_mod.f90:
// 81-127, riemann/*/riemann_mod.f90:41-101); here the rotation is an index
// permutation resolved at compile time, so no data moves.
//
// Real is the operator precision (double for the plain path, float for the
// FP32 operator).  Expressions keep the reference's evaluation order; the
// compiler may contract a*b+c into FMAs, which changes results by at most an
// ulp per operation.
#pragma once

#include <hip/hip_runtime.h>

namespace nemo {

enum { DENS = 0, MOMX = 1, MOMY = 2, MOMZ = 3, ENER = 4 };

// Momentum slot that plays the role of MOMX/MOMY/MOMZ (k = 0,1,2) when the
// problem is rotated to direction D.
template <int D, int K>
struct Mom {
    static constexpr int idx = 1 + (D + K) % 3;
};

template <typename Real>
struct Gas {
    Real kappa;        // ratio of specific heats
    Real kappa_m1;     // kappa - 1
    Real hllc_beta;    // sqrt(0.5*(kappa-1)/kappa)
    Real half_inv_km1; // 0.5/(kappa-1)
};

// pressure_elemental: equations/euler/equations_mod.f90:131-140.
// Evaluated without FMA contraction and, for a directional call, with the
// momentum squares summed in the order of the state rotated to that
// direction (the reference rotates before calling pressure: riemann/*/
// riemann_mod.f90:41-101, equations_mod.f90:436-493).  Reason: the hybrid
// HLLC/HLLE flux blends with sqrt(|p_L - p_R|/(p_L + p_R)), which is not
// Lipschitz at p_L = p_R, so a one-ulp difference in p between two
// evaluation orders moves the flux by ~1e-8 relative on flows of uniform
// pressure.  Bit-identical pressures keep the blend identical too.
template <typename Real>
__device__ __forceinline__ Real pressure_ordered(const Gas<Real>& g, Real rho, Real m0, Real m1, Real m2, Real E) {
#pragma clang fp contract(off)
    return g.kappa_m1 * (E - Real(0.5) / rho * (m0 * m0 + m1 * m1 + m2 * m2));
}

// Unrotated (x-ordered) pressure: isvalid, node primitives, indicator.
template <typename Real>
__device__ __forceinline__ Real pressure(const Gas<Real>& g, const Real* u) {
    return pressure_ordered(g, u[DENS], u[MOMX], u[MOMY], u[MOMZ], u[ENER]);
}

// Pressure of the state rotated to direction D.
template <int D, typename Real>
__device__ __forceinline__ Real pressure_dir(const Gas<Real>& g, const Real* u) {
    return pressure_ordered(g, u[DENS], u[Mom<D, 0>::idx], u[Mom<D, 1>::idx], u[Mom<D, 2>::idx], u[ENER]);
}

// xflux: equations/euler/equations_mod.f90:64-79, rotated to direction D.
template <int D, typename Real>
__device__ __forceinline__ void flux(const Gas<Real>& g, const Real* u, Real* f) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    const Real p = pressure_dir<D>(g, u);
    const Real vn = u[MN] / u[DENS];
    f[DENS] = u[MN];
    f[MN] = vn * u[MN] + p;
    f[MT1] = vn * u[MT1];
    f[MT2] = vn * u[MT2];
    f[ENER] = vn * (u[ENER] + p);
}

// xlambdaMax_elemental + soundspeed: equations_mod.f90:288-297, 384-393.
template <int D, typename Real>
__device__ __forceinline__ Real lambda_max(const Gas<Real>& g, const Real* u) {
    constexpr int MN = Mom<D, 0>::idx;
    const Real p = pressure_dir<D>(g, u);
    return fabs(u[MN] / u[DENS]) + sqrt(g.kappa * p / u[DENS]);
}

// Rusanov / local Lax-Friedrichs: equations/euler/riemann/rusanov/riemann_mod.f90:7-24
template <int D, typename Real>
__device__ __forceinline__ void riemann_rusanov(const Gas<Real>& g, const Real* uL, const Real* uR, Real* sf) {
    Real FL[5], FR[5];
    flux<D>(g, uL, FL);
    flux<D>(g, uR, FR);
    const Real a = lambda_max<D>(g, uL), b = lambda_max<D>(g, uR);
    const Real lmax = a > b ? a : b;
#pragma unroll
    for (int h = 0; h < 5; ++h) sf[h] = Real(0.5) * (FL[h] + FR[h] + lmax * (uL[h] - uR[h]));
}

// Hybrid HLLC/HLLE with Einfeldt-Roe speeds bounded by beta*c and the
// pressure-jump "smooth" blend: equations/euler/riemann/hllc/riemann_mod.f90:7-104
template <int D, typename Real>
__device__ __forceinline__ void riemann_hllc(const Gas<Real>& g, const Real* U_L, const Real* U_R, Real* F) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    const Real kappa = g.kappa;
    const Real V_L = U_L[MN] / U_L[DENS];
    const Real V_R = U_R[MN] / U_R[DENS];
    const Real p_L = pressure_dir<D>(g, U_L);
    const Real p_R = pressure_dir<D>(g, U_R);
    const Real c_L = sqrt(kappa * p_L / U_L[DENS]);
    const Real c_R = sqrt(kappa * p_R / U_R[DENS]);
    const Real H_L = (U_L[ENER] + p_L) / U_L[DENS];
    const Real H_R = (U_R[ENER] + p_R) / U_R[DENS];
    const Real sL = sqrt(U_L[DENS]), sR = sqrt(U_R[DENS]);
    const Real HTilde = (sL * H_L + sR * H_R) / (sL + sR);
    const Real vTilde = (sL * V_L + sR * V_R) / (sL + sR);
    const Real aTilde = sqrt(g.kappa_m1 * (HTilde - Real(0.5) * (vTilde * vTilde)));
    Real S_L = vTilde - aTilde, t = V_L - g.hllc_beta * c_L;
    S_L = S_L < t ? S_L : t;
    S_L = S_L < Real(0) ? S_L : Real(0);
    Real S_R = vTilde + aTilde;
    t = V_R + g.hllc_beta * c_R;
    S_R = S_R > t ? S_R : t;
    S_R = S_R > Real(0) ? S_R : Real(0);
    if (S_L >= Real(0)) {
        flux<D>(g, U_L, F);
    } else if (S_R <= Real(0)) {
        flux<D>(g, U_R, F);
    } else {
        Real F_L[5], F_R[5];
        flux<D>(g, U_L, F_L);
        flux<D>(g, U_R, F_R);
        const Real smooth = sqrt(fabs(p_L - p_R) / (p_L + p_R));
        const Real inv = Real(1) / (S_R - S_L);
        const Real SStar = (p_R - p_L + U_L[MN] * (S_L - V_L) - U_R[MN] * (S_R - V_R)) /
                           (U_L[DENS] * (S_L - V_L) - U_R[DENS] * (S_R - V_R));
        const bool left = (S_L <= Real(0)) && (Real(0) <= SStar);
        const Real* Us = left ? U_L : U_R;
        const Real* Fs = left ? F_L : F_R;
        const Real Ss = left ? S_L : S_R;
        const Real Vs = left ? V_L : V_R;
        const Real ps = left ? p_L : p_R;
        const Real QStar = Us[DENS] * (Ss - Vs) / (Ss - SStar);
        Real UStar[5];
        UStar[DENS] = QStar;
        UStar[MN] = QStar * SStar;
        UStar[MT1] = QStar * Us[MT1] / Us[DENS];
        UStar[MT2] = QStar * Us[MT2] / Us[DENS];
        UStar[ENER] = QStar * (Us[ENER] / Us[DENS] + (SStar - Vs) * (SStar + ps / Us[DENS] / (Ss - Vs)));
        const Real SLSR = S_L * S_R;
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            const Real hlle = (S_R * F_L[h] - S_L * F_R[h] + SLSR * (U_R[h] - U_L[h])) * inv;
            F[h] = smooth * hlle + (Real(1) - smooth) * (Fs[h] + Ss * (UStar[h] - Us[h]));
        }
    }
}

// logmean: equations/euler/two_point/chandrashekar/two_point_flux_mod.f90:13-32
// (series branch restated as written, including the sign of its u^2 term).
template <typename Real>
__device__ __forceinline__ Real logmean(Real a, Real b) {
    const Real x = a / b;
    if (fabs(a - b) < Real(1e-4)) {
        const Real c1 = Real(1.0 / 6.0), c2 = Real(2.0 / 45.0), c3 = Real(22.0 / 945.0);
        const Real u = (x * (x - Real(2)) + Real(1)) / (x * (x + Real(2)) + Real(1));
        return (a + b) * (Real(0.5) - u * (c1 - u * (c2 - c3 * u)));
    }
    return (a - b) / log(x);
}

// Per-node quantities of the two-point flux that depend on one state only
// (velocity, pressure, beta); computed once per node instead of once per
// pair.  Same expressions as two_point_flux_mod.f90:49-61.
template <typename Real>
struct NodePrim {
    Real rho, v[3], p, beta;
};

template <typename Real>
__device__ __forceinline__ NodePrim<Real> node_prim(const Gas<Real>& g, const Real* u) {
    NodePrim<Real> q;
    q.rho = u[DENS];
    q.p = pressure(g, u);
    q.beta = Real(0.5) * u[DENS] / q.p;
    q.v[0] = u[MOMX] / u[DENS];
    q.v[1] = u[MOMY] / u[DENS];
    q.v[2] = u[MOMZ] / u[DENS];
    return q;
}

// Chandrashekar entropy-conserving, kinetic-energy-preserving flux:
// two_point_flux_mod.f90:34-77, rotated to direction D (:79-141).
template <int D, typename Real>
__device__ __forceinline__ void ec_flux(const Gas<Real>& g, const NodePrim<Real>& a, const NodePrim<Real>& b, Real* f) {
    constexpr int N0 = D, N1 = (D + 1) % 3, N2 = (D + 2) % 3;
    const Real dens = logmean(a.rho, b.rho);
    const Real velx = Real(0.5) * (a.v[N0] + b.v[N0]);
    const Real vely = Real(0.5) * (a.v[N1] + b.v[N1]);
    const Real velz = Real(0.5) * (a.v[N2] + b.v[N2]);
    const Real pres = Real(0.5) * (a.rho + b.rho) / (a.beta + b.beta);
    const Real ssum = (a.v[N0] * a.v[N0] + b.v[N0] * b.v[N0]) + (a.v[N1] * a.v[N1] + b.v[N1] * b.v[N1]) +
                      (a.v[N2] * a.v[N2] + b.v[N2] * b.v[N2]);
    const Real enth = g.half_inv_km1 / logmean(a.beta, b.beta) + pres / dens + velx * velx + vely * vely +
                      velz * velz - Real(0.25) * ssum;
    const Real fr = dens * velx;
    f[DENS] = fr;
    f[1 + N0] = fr * velx + pres;
    f[1 + N1] = fr * vely;
    f[1 + N2] = fr * velz;
    f[ENER] = fr * enth;
}

// Central two-point flux: equations/euler/two_point/standard/two_point_flux_mod.f90:13-23
template <int D, typename Real>
__device__ __forceinline__ void central_flux(const Gas<Real>& g, const Real* u, const Real* v, Real* r) {
    Real fu[5], fv[5];
    flux<D>(g, u, fu);
    flux<D>(g, v, fv);
#pragma unroll
    for (int h = 0; h < 5; ++h) r[h] = Real(0.5) * (fu[h] + fv[h]);
}

// isvalid_elemental: equations_mod.f90:337-344
template <typename Real>
__device__ __forceinline__ bool isvalid(const Gas<Real>& g, const Real* u) {
    return (u[DENS] > Real(0)) && (pressure(g, u) > Real(0));
}

}  // namespace nemo
