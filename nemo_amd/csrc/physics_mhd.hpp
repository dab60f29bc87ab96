// Device physics for ideal GLM-MHD (gfx950): 9 variables
// (rho, rho v[3], E, B[3], psi), E = p/(kappa-1) + rho|v|^2/2 + (|B|^2 + psi^2)/(2 mu0).
//
// Each function restates one reference function (paths relative to the
// reference's source/ directory).  As for Euler, directions are compile-time
// index permutations of momentum and magnetic field together
// (equations/mhd/equations_mod.f90:141-220).  The GLM cleaning speed c_h and
// damping alpha (equations_share_mod glm_ch / glm_alpha) come from the CFL
// reduction of the step and are passed in `MhdGas::ch`.
#pragma once

#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace nemo {

enum { MAGX = 5, MAGY = 6, MAGZ = 7, GLMP = 8 };

template <typename Real>
struct MhdGas {
    Real kappa, kappa_m1, kappa_m2;
    Real mu0;
    Real smu0;         // 1/mu0
    Real half_inv_km1; // 0.5/(kappa-1)
    Real ch;           // GLM cleaning speed of the current step
};

// Momentum / field slot playing the role of x, y, z (k = 0,1,2) in direction D.
template <int D, int K>
struct Mag {
    static constexpr int idx = 5 + (D + K) % 3;
};

// pressure_elemental: equations/mhd/equations_mod.f90:223-234 (gas pressure,
// psi^2/2 taken out of the energy as well), in the reference's order and
// without FMA contraction: the CFL reduction is compared bit for bit.
template <typename Real>
__device__ __forceinline__ Real mhd_pressure(const MhdGas<Real>& g, const Real* u) {
#pragma clang fp contract(off)
    return g.kappa_m1 * (u[ENER] - div_ieee(Real(0.5), u[DENS]) * (u[MOMX] * u[MOMX] + u[MOMY] * u[MOMY] + u[MOMZ] * u[MOMZ]) -
                         div_ieee(Real(0.5), g.mu0) * (u[MAGX] * u[MAGX] + u[MAGY] * u[MAGY] + u[MAGZ] * u[MAGZ] + u[GLMP] * u[GLMP]));
}

template <typename Real>
__device__ __forceinline__ bool mhd_isvalid(const MhdGas<Real>& g, const Real* u) {
    return (u[DENS] > Real(0)) && (mhd_pressure(g, u) > Real(0));
}

// One node's state for a direction pass (computed once per node, used by
// every flux of the pass).
template <typename Real>
struct MhdState {
    Real u[9];
    Real inv_rho;
    Real v[3];
    Real p;      // gas pressure
    Real v2, B2; // |v|^2, |B|^2
    Real beta;   // 0.5 rho/p (EC flux only)
};

template <bool BETA, typename Real>
__device__ __forceinline__ void mhd_make_state(const MhdGas<Real>& g, MhdState<Real>& s) {
    s.inv_rho = frcp(s.u[DENS]);
    s.v[0] = s.u[MOMX] * s.inv_rho;
    s.v[1] = s.u[MOMY] * s.inv_rho;
    s.v[2] = s.u[MOMZ] * s.inv_rho;
    s.v2 = s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2];
    s.B2 = s.u[MAGX] * s.u[MAGX] + s.u[MAGY] * s.u[MAGY] + s.u[MAGZ] * s.u[MAGZ];
    s.p = g.kappa_m1 * (s.u[ENER] - Real(0.5) * (s.u[DENS] * s.v2 + g.smu0 * (s.B2 + s.u[GLMP] * s.u[GLMP])));
    if constexpr (BETA) s.beta = fdiv(Real(0.5) * s.u[DENS], s.p);
}

// xflux with the GLM terms: equations/mhd/equations_mod.f90:64-138, rotated to D.
template <int D, typename Real>
__device__ __forceinline__ void mhd_flux_s(const MhdGas<Real>& g, const MhdState<Real>& s, Real* f) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    constexpr int BN = Mag<D, 0>::idx, BT1 = Mag<D, 1>::idx, BT2 = Mag<D, 2>::idx;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    const Real psi = s.u[GLMP];
    const Real Etot = s.u[ENER] - Real(0.5) * g.smu0 * (psi * psi);
    const Real b1 = s.u[BN];
    const Real vn = s.v[D];
    const Real vb = s.u[MAGX] * s.v[0] + s.u[MAGY] * s.v[1] + s.u[MAGZ] * s.v[2];
    const Real pt = g.kappa_m1 * (Etot - Real(0.5) * s.u[DENS] * s.v2) - g.kappa_m2 * Real(0.5) * g.smu0 * s.B2;
    f[DENS] = s.u[MN];
    f[MN] = s.u[MN] * vn + pt - g.smu0 * b1 * b1;
    f[MT1] = s.u[MN] * s.v[T1] - g.smu0 * b1 * s.u[BT1];
    f[MT2] = s.u[MN] * s.v[T2] - g.smu0 * b1 * s.u[BT2];
    f[ENER] = (Etot + pt) * vn - g.smu0 * b1 * vb + g.smu0 * g.ch * b1 * psi;
    f[BN] = g.ch * psi;
    f[BT1] = vn * s.u[BT1] - b1 * s.v[T1];
    f[BT2] = vn * s.u[BT2] - b1 * s.v[T2];
    f[GLMP] = g.ch * b1;
}

// Fastest_Signal_Speed (equations/mhd/equations_mod.f90:693-714) in direction D.
// EXACT: IEEE square roots (the CFL reduction, compared bit for bit).
template <int D, bool EXACT = false, typename Real>
__device__ __forceinline__ Real mhd_fastspeed(const MhdGas<Real>& g, const MhdState<Real>& s) {
    constexpr int BN = Mag<D, 0>::idx, BT1 = Mag<D, 1>::idx, BT2 = Mag<D, 2>::idx;
    if constexpr (EXACT) {
#pragma clang fp contract(off)
        // the reference's expression order, IEEE divisions and roots
        const Real rho = s.u[DENS], mu0 = g.mu0;
        const Real cs2 = g.kappa * s.p / rho;
        const Real ca2 = s.u[BN] * s.u[BN] / rho / mu0;
        const Real va2 = (s.u[BT1] * s.u[BT1] + s.u[BT2] * s.u[BT2]) / rho / mu0 + ca2;
        const Real astar = sqrt((cs2 + va2) * (cs2 + va2) - Real(4) * cs2 * ca2);
        return sqrt(Real(0.5) * (cs2 + va2 + astar));
    } else {
        const Real cs2 = g.kappa * s.p * s.inv_rho;
        const Real ca2 = s.u[BN] * s.u[BN] * s.inv_rho * g.smu0;
        const Real va2 = (s.u[BT1] * s.u[BT1] + s.u[BT2] * s.u[BT2]) * s.inv_rho * g.smu0 + ca2;
        const Real astar = fsqrt((cs2 + va2) * (cs2 + va2) - Real(4) * cs2 * ca2);
        return fsqrt(Real(0.5) * (cs2 + va2 + astar));
    }
}

// x/y/zlambdaMax (equations/mhd/equations_mod.f90:522-690) from the raw
// state, FP64 (CFL reduction).
template <int D>
__device__ __forceinline__ double mhd_lambda_max(const MhdGas<double>& g, const double* u) {
    MhdState<double> s;
#pragma unroll
    for (int h = 0; h < 9; ++h) s.u[h] = u[h];
    s.inv_rho = 1.0 / u[DENS];
    s.v[0] = u[MOMX] / u[DENS];
    s.v[1] = u[MOMY] / u[DENS];
    s.v[2] = u[MOMZ] / u[DENS];
    s.p = mhd_pressure(g, u);
    return fabs(s.v[D]) + mhd_fastspeed<D, true>(g, s);
}

// Rusanov: equations/mhd/riemann/rusanov/riemann_mod.f90:7-24
template <int D, typename Real>
__device__ __forceinline__ void mhd_rusanov(const MhdGas<Real>& g, const MhdState<Real>& L, const MhdState<Real>& R,
                                            Real* sf) {
    Real FL[9], FR[9];
    mhd_flux_s<D>(g, L, FL);
    mhd_flux_s<D>(g, R, FR);
    const Real a = fabs(L.v[D]) + mhd_fastspeed<D>(g, L);
    const Real b = fabs(R.v[D]) + mhd_fastspeed<D>(g, R);
    const Real lmax = a > b ? a : b;
#pragma unroll
    for (int h = 0; h < 9; ++h) sf[h] = Real(0.5) * (FL[h] + FR[h] + lmax * (L.u[h] - R.u[h]));
}

// Powell + GLM non-conservative terms (equations/mhd/non_conservative_flux_mod.f90:
// 10-45): Phi(u_I) times the interface averages {{B_n}} and {{psi}}.  The
// averages are shared by both nodes of a pair, so the caller forms them once.
template <int D, typename Real>
__device__ __forceinline__ void mhd_nc_add(const MhdGas<Real>& g, const MhdState<Real>& I, Real Bavg, Real Gavg,
                                           Real* f) {
    const Real vb = I.u[MAGX] * I.v[0] + I.u[MAGY] * I.v[1] + I.u[MAGZ] * I.v[2];
    const Real sB = Bavg * g.smu0;
    f[MOMX] += sB * I.u[MAGX];
    f[MOMY] += sB * I.u[MAGY];
    f[MOMZ] += sB * I.u[MAGZ];
    f[ENER] += sB * vb + Gavg * I.v[D] * (I.u[GLMP] * g.smu0);
    f[MAGX] += Bavg * I.v[0];
    f[MAGY] += Bavg * I.v[1];
    f[MAGZ] += Bavg * I.v[2];
    f[GLMP] += Gavg * I.v[D];
}

// LN_MEAN: equations/mhd/two_point/ec/two_point_flux_mod.f90:203-219, with
// its reciprocal.  FP32: the Ismail-Roe series form (logmean2_select).
// Both branches are evaluated and selected (straight-line code, as the
// Euler log-mean); Xi takes the IEEE-rounded bits (div_rn), which the log
// branch amplifies by 1/|Xi - 1|.
__device__ __forceinline__ void ln_mean(double aL, double aR, double& lm, double& inv_lm) {
    const double Xi = div_rn(aR, aL);
    const double u = (Xi * (Xi - 2.0) + 1.0) * rcp_nr(Xi * (Xi + 2.0) + 1.0);
    const bool series = u < 1.0e-4;
    const double d = 105.0 + u * (35.0 + u * (21.0 + u * 15.0));
    const double s = (aL + aR) * 52.5;
    const double l = log_pos(Xi);
    const double dif = aR - aL;
    // series: lm = s/d, 1/lm = d/s; log: lm = dif/l, 1/lm = l/dif
    const double num = series ? s : dif, den = series ? d : l;
    lm = num * rcp_nr(den);
    inv_lm = den * rcp_nr(num);
}
__device__ __forceinline__ void ln_mean(float aL, float aR, float& lm, float& inv_lm) {
    logmean2_select(aL, aR, lm, inv_lm);
}

// Entropy-conserving GLM-MHD flux of Derigs et al. (JCP 364, 2018):
// equations/mhd/two_point/ec/two_point_flux_mod.f90:13-86, rotated to D
// (the reference's own y/z versions do not compile).
template <int D, typename Real>
__device__ __forceinline__ void mhd_ec_flux(const MhdGas<Real>& g, const MhdState<Real>& L, const MhdState<Real>& R,
                                            Real* Fs) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    constexpr int BN = Mag<D, 0>::idx, BT1 = Mag<D, 1>::idx, BT2 = Mag<D, 2>::idx;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    Real rhoLN, inv_rhoLN, betaLN, inv_betaLN;
    ln_mean(L.u[DENS], R.u[DENS], rhoLN, inv_rhoLN);
    ln_mean(L.beta, R.beta, betaLN, inv_betaLN);
    const Real vn = Real(0.5) * (L.v[D] + R.v[D]);
    const Real vt1 = Real(0.5) * (L.v[T1] + R.v[T1]);
    const Real vt2 = Real(0.5) * (L.v[T2] + R.v[T2]);
    const Real bn = Real(0.5) * (L.u[BN] + R.u[BN]);
    const Real bt1 = Real(0.5) * (L.u[BT1] + R.u[BT1]);
    const Real bt2 = Real(0.5) * (L.u[BT2] + R.u[BT2]);
    const Real v1_B2Avg = Real(0.5) * (L.v[D] * L.B2 + R.v[D] * R.B2);
    const Real vbL = L.v[0] * L.u[MAGX] + L.v[1] * L.u[MAGY] + L.v[2] * L.u[MAGZ];
    const Real vbR = R.v[0] * R.u[MAGX] + R.v[1] * R.u[MAGY] + R.v[2] * R.u[MAGZ];
    const Real vB_Avg = Real(0.5) * (vbL + vbR);
    const Real pTilde = Real(0.5) * (fdiv(L.u[DENS] + R.u[DENS], L.beta + R.beta) + g.smu0 * Real(0.5) * (L.B2 + R.B2));
    const Real psiAvg = Real(0.5) * (L.u[GLMP] + R.u[GLMP]);
    const Real f0 = rhoLN * vn;
    Fs[DENS] = f0;
    Fs[MN] = f0 * vn - g.smu0 * bn * bn + pTilde;
    Fs[MT1] = f0 * vt1 - g.smu0 * bn * bt1;
    Fs[MT2] = f0 * vt2 - g.smu0 * bn * bt2;
    Fs[BT1] = vn * bt1 - bn * vt1;
    Fs[BT2] = vn * bt2 - bn * vt2;
    Fs[BN] = g.ch * psiAvg;
    Fs[GLMP] = g.ch * bn;
    Fs[ENER] = f0 * Real(0.5) * (Real(2) * g.half_inv_km1 * inv_betaLN - Real(0.5) * (L.v2 + R.v2)) +
               (vn * Fs[MN] + vt1 * Fs[MT1] + vt2 * Fs[MT2]) +
               g.smu0 * ((bn * Fs[BN] + bt1 * Fs[BT1] + bt2 * Fs[BT2]) - Real(0.5) * v1_B2Avg + bn * vB_Avg +
                         Fs[GLMP] * psiAvg - g.ch * Real(0.5) * (L.u[GLMP] * L.u[BN] + R.u[GLMP] * R.u[BN]));
}

// Central two-point flux (equations/mhd/two_point/standard): mean of the
// physical fluxes.
template <int D, typename Real>
__device__ __forceinline__ void mhd_central_flux(const MhdGas<Real>& g, const MhdState<Real>& a,
                                                 const MhdState<Real>& b, Real* r) {
    Real fu[9], fv[9];
    mhd_flux_s<D>(g, a, fu);
    mhd_flux_s<D>(g, b, fv);
#pragma unroll
    for (int h = 0; h < 9; ++h) r[h] = Real(0.5) * (fu[h] + fv[h]);
}

// HLLD (Miyoshi & Kusano, JCP 208, 2005) after
// equations/mhd/riemann/hlld/riemann_mod.f90:16-248, whose shipped form does
// not compile (an "# Error" guard) and has no GLM coupling.  Two changes,
// the same as the FP64 oracle's: fast speeds from the conservative states,
// and the (B_n, psi) subsystem solved exactly first,
//   B_n* = {{B_n}} + (psi_L - psi_R)/2,  psi* = {{psi}} + (B_nL - B_nR)/2,
// F(B_n) = c_h psi*, F(psi) = c_h B_n*, energy + c_h B_n* psi*/mu0, with the
// seven-wave solver run on B_n = B_n* and the energy without psi^2/2.
// Works on the state rotated to direction D (r[] in x-order), writes F in
// the unrotated variable order.
template <int D, typename Real>
__device__ __forceinline__ void mhd_hlld(const MhdGas<Real>& g, const MhdState<Real>& SL_, const MhdState<Real>& SR_,
                                         Real* Fout) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    constexpr int BN = Mag<D, 0>::idx, BT1 = Mag<D, 1>::idx, BT2 = Mag<D, 2>::idx;
    constexpr int T1 = (D + 1) % 3, T2 = (D + 2) % 3;
    const Real hlld_eps = Real(1.0e-8), hlld_small_eps = Real(1.0e-12);
    const Real smu0 = g.smu0;
    const Real Bstar = Real(0.5) * (SL_.u[BN] + SR_.u[BN]) + Real(0.5) * (SL_.u[GLMP] - SR_.u[GLMP]);
    const Real psistar = Real(0.5) * (SL_.u[GLMP] + SR_.u[GLMP]) + Real(0.5) * (SL_.u[BN] - SR_.u[BN]);
    const Real Bn = Bstar;

    // primitive states in the rotated frame with B_n = B_n*:
    // rho, u[3] (normal first), p, B[3]
    Real rho[2], un[2], ut1[2], ut2[2], p[2], bt1[2], bt2[2], e[2], cons[2][8], F[2][8];
    {
        const MhdState<Real>* S[2] = {&SL_, &SR_};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const MhdState<Real>& X = *S[s];
            rho[s] = X.u[DENS];
            un[s] = X.v[D];
            ut1[s] = X.v[T1];
            ut2[s] = X.v[T2];
            p[s] = X.p;  // gas pressure (psi^2 and B^2 removed)
            bt1[s] = X.u[BT1];
            bt2[s] = X.u[BT2];
            const Real v2 = X.v2;
            const Real b2 = Bn * Bn + bt1[s] * bt1[s] + bt2[s] * bt2[s];
            e[s] = p[s] * frcp(g.kappa_m1) + Real(0.5) * rho[s] * v2 + Real(0.5) * smu0 * b2;
            // conservative state (rotated, without psi): rho, m[3], e, B[3]
            cons[s][0] = rho[s];
            cons[s][1] = rho[s] * un[s];
            cons[s][2] = rho[s] * ut1[s];
            cons[s][3] = rho[s] * ut2[s];
            cons[s][4] = e[s];
            cons[s][5] = Bn;
            cons[s][6] = bt1[s];
            cons[s][7] = bt2[s];
            // x flux without GLM (c_h = 0)
            const Real vb = Bn * un[s] + bt1[s] * ut1[s] + bt2[s] * ut2[s];
            const Real pt = g.kappa_m1 * (e[s] - Real(0.5) * rho[s] * v2) - g.kappa_m2 * Real(0.5) * smu0 * b2;
            F[s][0] = cons[s][1];
            F[s][1] = cons[s][1] * un[s] + pt - smu0 * Bn * Bn;
            F[s][2] = cons[s][1] * ut1[s] - smu0 * Bn * bt1[s];
            F[s][3] = cons[s][1] * ut2[s] - smu0 * Bn * bt2[s];
            F[s][4] = (e[s] + pt) * un[s] - smu0 * Bn * vb;
            F[s][5] = Real(0);
            F[s][6] = un[s] * bt1[s] - Bn * ut1[s];
            F[s][7] = un[s] * bt2[s] - Bn * ut2[s];
        }
    }
    Real cf[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const Real ir = frcp(rho[s]);
        const Real cs2 = g.kappa * p[s] * ir;
        const Real ca2 = Bn * Bn * ir * smu0;
        const Real va2 = (bt1[s] * bt1[s] + bt2[s] * bt2[s]) * ir * smu0 + ca2;
        const Real astar = fsqrt((cs2 + va2) * (cs2 + va2) - Real(4) * cs2 * ca2);
        cf[s] = fsqrt(Real(0.5) * (cs2 + va2 + astar));
    }
    const Real fw = cf[0] > cf[1] ? cf[0] : cf[1];
    const Real SL = (un[0] < un[1] ? un[0] : un[1]) - fw;
    const Real SR = (un[0] > un[1] ? un[0] : un[1]) + fw;
    Real Fl[8];
    if (SL > Real(0)) {
#pragma unroll
        for (int h = 0; h < 8; ++h) Fl[h] = F[0][h];
    } else if (SR < Real(0)) {
#pragma unroll
        for (int h = 0; h < 8; ++h) Fl[h] = F[1][h];
    } else {
        const Real pmagL = Real(0.5) * (Bn * Bn + bt1[0] * bt1[0] + bt2[0] * bt2[0]);
        const Real pmagR = Real(0.5) * (Bn * Bn + bt1[1] * bt1[1] + bt2[1] * bt2[1]);
        const Real ptotL = p[0] + pmagL, ptotR = p[1] + pmagR;
        const Real dL = rho[0] * (SL - un[0]), dR = rho[1] * (SR - un[1]);
        const Real SM = fdiv(dR * un[1] - dL * un[0] - ptotR + ptotL, dR - dL);
        const Real sSL_SM = frcp(SL - SM), sSR_SM = frcp(SR - SM);
        const Real rhoLS = dL * sSL_SM, rhoRS = dR * sSR_SM;
        const Real sqL = fsqrt(rhoLS), sqR = fsqrt(rhoRS);
        const Real SLS = SM - fabs(Bn) * frcp(sqL);
        const Real SRS = SM + fabs(Bn) * frcp(sqR);
        const Real ptotS = ptotL + dL * (SM - un[0]);
        // single-star states (tangential parts)
        Real utLS[2], btLS[2], utRS[2], btRS[2];
        {
            const Real temp = dL * (SL - SM) - Bn * Bn;
            if (fabs(temp) < hlld_eps) {
                utLS[0] = ut1[0]; utLS[1] = ut2[0]; btLS[0] = bt1[0]; btLS[1] = bt2[0];
            } else {
                const Real st = frcp(temp);
                const Real cu = Bn * (SM - un[0]) * st, cb = (dL * (SL - un[0]) - Bn * Bn) * st;
                utLS[0] = ut1[0] - cu * bt1[0];
                utLS[1] = ut2[0] - cu * bt2[0];
                btLS[0] = cb * bt1[0];
                btLS[1] = cb * bt2[0];
            }
        }
        {
            const Real temp = dR * (SR - SM) - Bn * Bn;
            if (fabs(temp) < hlld_eps) {
                utRS[0] = ut1[1]; utRS[1] = ut2[1]; btRS[0] = bt1[1]; btRS[1] = bt2[1];
            } else {
                const Real st = frcp(temp);
                const Real cu = Bn * (SM - un[1]) * st, cb = (dR * (SR - un[1]) - Bn * Bn) * st;
                utRS[0] = ut1[1] - cu * bt1[1];
                utRS[1] = ut2[1] - cu * bt2[1];
                btRS[0] = cb * bt1[1];
                btRS[1] = cb * bt2[1];
            }
        }
        const Real uBL = un[0] * Bn + ut1[0] * bt1[0] + ut2[0] * bt2[0];
        const Real uBR = un[1] * Bn + ut1[1] * bt1[1] + ut2[1] * bt2[1];
        const Real uBLS = SM * Bn + utLS[0] * btLS[0] + utLS[1] * btLS[1];
        const Real uBRS = SM * Bn + utRS[0] * btRS[0] + utRS[1] * btRS[1];
        const Real eLS = (e[0] * (SL - un[0]) + ptotS * SM - ptotL * un[0] + Bn * (uBL - uBLS)) * sSL_SM;
        const Real eRS = (e[1] * (SR - un[1]) - ptotR * un[1] + ptotS * SM + Bn * (uBR - uBRS)) * sSR_SM;
        Real ULS[8] = {rhoLS, rhoLS * SM, rhoLS * utLS[0], rhoLS * utLS[1], eLS, Bn, btLS[0], btLS[1]};
        Real URS[8] = {rhoRS, rhoRS * SM, rhoRS * utRS[0], rhoRS * utRS[1], eRS, Bn, btRS[0], btRS[1]};
        Real ULSS[8], URSS[8];
        if (fabs(Bn) < hlld_small_eps) {
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                ULSS[h] = ULS[h];
                URSS[h] = URS[h];
            }
        } else {
            const Real sgn = Bn < Real(0) ? Real(-1) : Real(1);
            const Real st = frcp(sqL + sqR);
            Real utSS[2], btSS[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                utSS[k] = (sqL * utLS[k] + sqR * utRS[k] + sgn * (btRS[k] - btLS[k])) * st;
                btSS[k] = (sqL * btRS[k] + sqR * btLS[k] + sgn * (utRS[k] - utLS[k]) * sqL * sqR) * st;
            }
            // the reference's tangential products (normal parts vanish)
            const Real uBSS = utSS[0] * btSS[0] + utSS[1] * btSS[1];
            const Real uBL2 = utLS[0] * btLS[0] + utLS[1] * btLS[1];
            const Real uBR2 = utRS[0] * btRS[0] + utRS[1] * btRS[1];
            const Real eLSS = eLS - sgn * sqL * (uBL2 - uBSS);
            const Real eRSS = eRS + sgn * sqR * (uBR2 - uBSS);
            ULSS[0] = rhoLS; URSS[0] = rhoRS;
            ULSS[1] = rhoLS * SM; URSS[1] = rhoRS * SM;
            ULSS[2] = rhoLS * utSS[0]; URSS[2] = rhoRS * utSS[0];
            ULSS[3] = rhoLS * utSS[1]; URSS[3] = rhoRS * utSS[1];
            ULSS[4] = eLSS; URSS[4] = eRSS;
            ULSS[5] = Bn; URSS[5] = Bn;
            ULSS[6] = btSS[0]; URSS[6] = btSS[0];
            ULSS[7] = btSS[1]; URSS[7] = btSS[1];
        }
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            if (Real(0) <= SLS)
                Fl[h] = F[0][h] + SL * (ULS[h] - cons[0][h]);
            else if (Real(0) <= SM)
                Fl[h] = F[0][h] + SLS * ULSS[h] - (SLS - SL) * ULS[h] - SL * cons[0][h];
            else if (Real(0) <= SRS)
                Fl[h] = F[1][h] + SRS * URSS[h] - (SRS - SR) * URS[h] - SR * cons[1][h];
            else
                Fl[h] = F[1][h] + SR * (URS[h] - cons[1][h]);
        }
    }
    // GLM subsystem and rotation back
    Fout[DENS] = Fl[0];
    Fout[MN] = Fl[1];
    Fout[MT1] = Fl[2];
    Fout[MT2] = Fl[3];
    Fout[ENER] = Fl[4] + g.ch * Bstar * psistar * smu0;
    Fout[BN] = g.ch * psistar;
    Fout[BT1] = Fl[6];
    Fout[BT2] = Fl[7];
    Fout[GLMP] = g.ch * Bstar;
}

}  // namespace nemo
