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

// Arithmetic by operator precision.  Where the reference's result is not
// sensitive to the last bit, divisions and reciprocals use the hardware
// reciprocal estimate: v_rcp_f32 for FP32 (about 1 ulp), and for FP64
// v_rcp_f64 (relative error below 2^-24) refined by one third-order step,
// r (1 + e + e^2) with e = 1 - b r, whose error is e^3 before the final
// rounding: 4 instructions, against 5 for two Newton steps and 10 for an
// IEEE division; on gfx950 it returned the correctly rounded reciprocal for
// all but 1 of 6.7e7 sampled arguments (nemo_gpu_check_arithmetic).
// IEEE-rounded division stays where bits matter (div_rn below): the
// pressure (see pressure_ordered), the ratio a/b
// of the log-mean, beta and the CFL speeds.
__device__ __forceinline__ double rcp_nr(double b) {
    const double r = __builtin_amdgcn_rcp(b);
    const double e = fma(-b, r, 1.0);
    return fma(r, fma(e, e, e), r);
}
// fma(a, b, c) for a constant c: the compiler otherwise copies every 64-bit
// constant into a vector register pair (two v_mov per use, twenty per log
// below) before a v_fmac; v_fma_f64 reads c from a scalar register pair,
// set by s_mov on the scalar unit.  Same fused operation, same bits.
__device__ __forceinline__ double fma_sc(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}
// Correctly rounded a / b, the bits of the IEEE division, for operands that
// keep b, 1/b and a/b normal and far from overflow (here: densities,
// pressures and their ratios): y = 1/b (rcp_nr), q = a y, and one
// Markstein correction q + (a - b q) y (P. Markstein, IBM J. Res. Dev. 34,
// 1990).  These are the correction steps of the compiler's division; its
// v_div_scale range scaling, v_div_fixup special cases and the VCC-carried
// v_div_fmas are left out, so independent divisions overlap.
__device__ __forceinline__ double div_rn(double a, double b) {
    const double y = rcp_nr(b);
    const double q = a * y;
    return fma(fma(-b, q, a), y, q);
}
// IEEE-rounded division: div_rn in FP64 (same bits for the operands here),
// the IEEE division in FP32.
__device__ __forceinline__ double div_ieee(double a, double b) { return div_rn(a, b); }
__device__ __forceinline__ float div_ieee(float a, float b) { return a / b; }
__device__ __forceinline__ double fdiv(double a, double b) { return a * rcp_nr(b); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double frcp(double a) { return rcp_nr(a); }
__device__ __forceinline__ float frcp(float a) { return __builtin_amdgcn_rcpf(a); }
// Square root to about 1 ulp for a >= 0: v_rsq_f64 estimate, one Goldschmidt
// step and a Newton correction (8 instructions against about 15 for the
// IEEE-rounded sqrt with its range scaling; arguments here are O(1e-12..1e12)).
__device__ __forceinline__ double fsqrt(double a) {
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, a);
    g = fma(d, h, g);
    return a > 0.0 ? g : 0.0;
}
__device__ __forceinline__ float fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }

// Natural logarithm of a positive finite x to about 2 ulp (the library log
// carries double-double arithmetic for 0.5 ulp, three times the work):
// x = m 2^k with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
// s = (m-1)/(m+1), |s| <= 0.1716, summed as 2s + 2s^3 sum_j s^2j/(2j+3) up to
// s^21 (truncation below 1e-18 relative; Abramowitz & Stegun 4.1.27), plus
// k log 2.  m - 1 is exact, so the result keeps its relative accuracy for x
// near 1, where the log-mean divides by it.  The three-argument form also
// returns t = s^2; with unscaled set it takes m = x, k = 0 (then t is
// ((x-1)/(x+1))^2 for any x, and the log is valid only for x near 1).
// log_series is the reduced part, log m for m in [sqrt(1/2), sqrt(2)), with t.
__device__ __forceinline__ double log_series(double m, double& t_out) {
    const double f = m - 1.0;
    const double s = f * rcp_nr(m + 1.0);
    const double t = s * s;
    double P;
    {
        // P = 1/3 + t/5 + ... + t^8/19 + t^9/21 in Horner form; its first
        // step is unfused because two constants cannot share one fma (one
        // scalar operand per instruction).  The dropped t^10/23 term is
        // below 1e-18 relative.
#pragma clang fp contract(off)
        P = t * (1.0 / 21.0) + 1.0 / 19.0;
    }
    P = fma_sc(P, t, 1.0 / 17.0);
    P = fma_sc(P, t, 1.0 / 15.0);
    P = fma_sc(P, t, 1.0 / 13.0);
    P = fma_sc(P, t, 1.0 / 11.0);
    P = fma_sc(P, t, 1.0 / 9.0);
    P = fma_sc(P, t, 1.0 / 7.0);
    P = fma_sc(P, t, 1.0 / 5.0);
    P = fma_sc(P, t, 1.0 / 3.0);
    t_out = t;
    return fma(2.0 * s * t, P, 2.0 * s);
}
__device__ __forceinline__ double log_pos(double x, bool unscaled, double& t_out) {
    int k = __builtin_amdgcn_frexp_exp(x);
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = m + m;
        k = k - 1;
    }
    if (unscaled) {
        m = x;
        k = 0;
    }
    const double lm = log_series(m, t_out);
    return fma(double(k), 0.69314718055994530942, lm);
}
__device__ __forceinline__ double log_pos(double x) {
    double t;
    return log_pos(x, false, t);
}
__device__ __forceinline__ double flog(double a) { return log_pos(a); }
__device__ __forceinline__ float flog(float a) { return __builtin_amdgcn_logf(a) * 0.69314718055994530942f; }

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
    return g.kappa_m1 * (E - div_ieee(Real(0.5), rho) * (m0 * m0 + m1 * m1 + m2 * m2));
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
    return fabs(div_ieee(u[MN], u[DENS])) + sqrt(div_ieee(g.kappa * p, u[DENS]));
}

// One node's state as seen from direction D: everything the two-point flux
// and the Riemann solvers need, computed once per node and direction pass
// instead of once per flux evaluation.  The reference recomputes these
// quantities inside every flux (e.g. two_point_flux_mod.f90:49-61,
// riemann/hllc/riemann_mod.f90:36-56).  Divisions by rho become products
// with 1/rho, which moves results by at most an ulp; the pressure keeps the
// reference's exact expression and rotated summation order (see
// pressure_ordered) because the HLLC/HLLE blend is not Lipschitz in p_L-p_R.
template <typename Real>
struct NodeState {
    Real u[5];
    Real inv_rho;
    Real v[3];   // velocity components x, y, z
    Real p;      // pressure, rotated order of the pass direction
    Real beta;   // 0.5*rho/p (two-point flux only)
    double p64;  // FP32 operator only: the FP64 pressure, for the HLLC/HLLE blend
};

template <int D, bool BETA, typename Real>
__device__ __forceinline__ void make_state(const Gas<Real>& g, NodeState<Real>& s) {
    s.inv_rho = frcp(s.u[DENS]);
    s.v[0] = s.u[MOMX] * s.inv_rho;
    s.v[1] = s.u[MOMY] * s.inv_rho;
    s.v[2] = s.u[MOMZ] * s.inv_rho;
    if constexpr (sizeof(Real) == 8) {
        s.p = pressure_dir<D>(g, s.u);
    } else {
        s.p = g.kappa_m1 * (s.u[ENER] - Real(0.5) * s.inv_rho *
                                            (s.u[MOMX] * s.u[MOMX] + s.u[MOMY] * s.u[MOMY] + s.u[MOMZ] * s.u[MOMZ]));
    }
    // beta enters the log-mean as the ratio beta_a / beta_b, whose rounding
    // the log branch amplifies by 1 / |beta_a / beta_b - 1| (up to 1e4 at the
    // series switch): in FP64 it takes the reference's IEEE-rounded bits.
    if constexpr (BETA) {
        if constexpr (sizeof(Real) == 8)
            s.beta = div_rn(Real(0.5) * s.u[DENS], s.p);
        else
            s.beta = fdiv(Real(0.5) * s.u[DENS], s.p);
    }
}

// xflux (equations_mod.f90:64-79) from a node state.
template <int D, typename Real>
__device__ __forceinline__ void flux_s(const NodeState<Real>& s, Real* f) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    const Real vn = s.v[D];
    f[DENS] = s.u[MN];
    f[MN] = vn * s.u[MN] + s.p;
    f[MT1] = vn * s.u[MT1];
    f[MT2] = vn * s.u[MT2];
    f[ENER] = vn * (s.u[ENER] + s.p);
}

// Rusanov / local Lax-Friedrichs: equations/euler/riemann/rusanov/riemann_mod.f90:7-24
template <int D, typename Real>
__device__ __forceinline__ void riemann_rusanov(const Gas<Real>& g, const NodeState<Real>& L, const NodeState<Real>& R,
                                                Real* sf) {
    Real FL[5], FR[5];
    flux_s<D>(L, FL);
    flux_s<D>(R, FR);
    const Real a = fabs(L.v[D]) + fsqrt(g.kappa * L.p * L.inv_rho);
    const Real b = fabs(R.v[D]) + fsqrt(g.kappa * R.p * R.inv_rho);
    const Real lmax = a > b ? a : b;
#pragma unroll
    for (int h = 0; h < 5; ++h) sf[h] = Real(0.5) * (FL[h] + FR[h] + lmax * (L.u[h] - R.u[h]));
}

// Hybrid HLLC/HLLE with Einfeldt-Roe speeds bounded by beta*c and the
// pressure-jump "smooth" blend: equations/euler/riemann/hllc/riemann_mod.f90:7-104
template <int D, typename Real>
__device__ __forceinline__ void riemann_hllc(const Gas<Real>& g, const NodeState<Real>& L, const NodeState<Real>& R,
                                             Real* F) {
    constexpr int MN = Mom<D, 0>::idx, MT1 = Mom<D, 1>::idx, MT2 = Mom<D, 2>::idx;
    const Real V_L = L.v[D], V_R = R.v[D];
    const Real p_L = L.p, p_R = R.p;
    const Real c_L = fsqrt(g.kappa * p_L * L.inv_rho);
    const Real c_R = fsqrt(g.kappa * p_R * R.inv_rho);
    const Real H_L = (L.u[ENER] + p_L) * L.inv_rho;
    const Real H_R = (R.u[ENER] + p_R) * R.inv_rho;
    const Real sL = fsqrt(L.u[DENS]), sR = fsqrt(R.u[DENS]);
    const Real inv_s = frcp(sL + sR);
    const Real HTilde = (sL * H_L + sR * H_R) * inv_s;
    const Real vTilde = (sL * V_L + sR * V_R) * inv_s;
    const Real aTilde = fsqrt(g.kappa_m1 * (HTilde - Real(0.5) * (vTilde * vTilde)));
    Real S_L = vTilde - aTilde, t = V_L - g.hllc_beta * c_L;
    S_L = S_L < t ? S_L : t;
    S_L = S_L < Real(0) ? S_L : Real(0);
    Real S_R = vTilde + aTilde;
    t = V_R + g.hllc_beta * c_R;
    S_R = S_R > t ? S_R : t;
    S_R = S_R > Real(0) ? S_R : Real(0);
    if (S_L >= Real(0)) {
        flux_s<D>(L, F);
    } else if (S_R <= Real(0)) {
        flux_s<D>(R, F);
    } else {
        Real F_L[5], F_R[5];
        flux_s<D>(L, F_L);
        flux_s<D>(R, F_R);
        Real smooth;
        if constexpr (sizeof(Real) == 8) {
            smooth = fsqrt(fabs(p_L - p_R) * frcp(p_L + p_R));
        } else {
            // The blend weight sqrt(|dp|/sum p) turns a relative pressure
            // error e into a weight of about sqrt(e): FP32 pressures would
            // give ~2e-4 instead of the FP64 ~1e-8 on uniform-pressure
            // contacts, so the FP32 operator takes dp from FP64 pressures.
            smooth = fsqrt(float(fabs(L.p64 - R.p64)) * frcp(float(L.p64 + R.p64)));
        }
        const Real inv = frcp(S_R - S_L);
        const Real SStar = fdiv(p_R - p_L + L.u[MN] * (S_L - V_L) - R.u[MN] * (S_R - V_R),
                                L.u[DENS] * (S_L - V_L) - R.u[DENS] * (S_R - V_R));
        const bool left = (S_L <= Real(0)) && (Real(0) <= SStar);
        const NodeState<Real>& X = left ? L : R;
        const Real* Fs = left ? F_L : F_R;
        const Real Ss = left ? S_L : S_R;
        const Real Vs = X.v[D];
        const Real QStar = fdiv(X.u[DENS] * (Ss - Vs), Ss - SStar);
        Real UStar[5];
        UStar[DENS] = QStar;
        UStar[MN] = QStar * SStar;
        UStar[MT1] = QStar * X.v[MT1 - 1];
        UStar[MT2] = QStar * X.v[MT2 - 1];
        UStar[ENER] = QStar * (X.u[ENER] * X.inv_rho + (SStar - Vs) * (SStar + fdiv(X.p * X.inv_rho, Ss - Vs)));
        const Real SLSR = S_L * S_R;
#pragma unroll
        for (int h = 0; h < 5; ++h) {
            const Real hlle = (S_R * F_L[h] - S_L * F_R[h] + SLSR * (R.u[h] - L.u[h])) * inv;
            F[h] = smooth * hlle + (Real(1) - smooth) * (Fs[h] + Ss * (UStar[h] - X.u[h]));
        }
    }
}

// logmean, equations/euler/two_point/chandrashekar/two_point_flux_mod.f90:
// 13-32 (series branch as written, including the sign of its u^2 term),
// returned together with its reciprocal so that the flux needs no further
// division by it: 1/L = log(a/b)/(a-b) on the log branch.  NEAR1: the
// caller guarantees a/b in [sqrt(1/2), sqrt(2)) after rounding, where the
// log's range reduction leaves m = a/b and k = 0; it is skipped, the
// remaining operations are the same, and so are the bits.
template <bool NEAR1 = false>
__device__ __forceinline__ void logmean2(double a, double b, double& lm, double& inv_lm) {
    // Both branches are evaluated and selected: straight-line code lets the
    // scheduler interleave the independent log-means of a line, where
    // per-pair branches serialise them.
    const double x = div_rn(a, b);  // IEEE bits: log(x) amplifies its rounding by 1/|x-1|
    const double d = a - b;
    const bool series = fabs(d) < 1e-4;
    const double c1 = 1.0 / 6.0, c2 = 2.0 / 45.0, c3 = 22.0 / 945.0;
    // The series branch's u = (x(x-2)+1)/(x(x+2)+1) = ((x-1)/(x+1))^2 is the
    // t of the log's reduction taken without 2^k scaling, which the series
    // lanes select (their log is not used).  Where |a-b| < 1e-4 and a, b are
    // O(1), u < 1e-8 and enters only as u/6 against 1/2: the difference
    // between the two roundings of it is below 1e-16 of the log-mean.
    double u;
    const double l = NEAR1 ? log_series(x, u) : log_pos(x, series, u);
    double w;
    {
#pragma clang fp contract(off)
        w = c2 - c3 * u;
    }
    const double ls = (a + b) * (0.5 - u * fma_sc(-u, w, c1));
    const double r = rcp_nr(series ? ls : d);
    lm = series ? ls : d * rcp_nr(l);
    inv_lm = series ? r : l * r;
}

// FP32 log-mean.  The reference's log branch loses ~eps/|a/b-1| relative
// accuracy near its switch |a-b| = 1e-4 (harmless in FP64, 6e-4 in FP32), so
// the FP32 operator uses the series of Ismail & Roe (JCP 228, 2009) in
// f = (a-b)/(a+b):  L = (a+b) / (2 sum_k f^2k/(2k+1)), ten terms for
// f^2 < 0.2 (ratios a/b up to 2.6; truncation below 5e-9), and the log form
// beyond, where |log(a/b)| > 0.9.  The log form is evaluated only when some
// lane of the wavefront needs it (a wavefront-uniform branch): in smooth
// flow a pair then costs one reciprocal (two with the mean itself) instead
// of five transcendental instructions, which issue at a quarter of the
// VALU rate.  (NEAR1 has no use here: the series covers those ratios.)
template <bool NEAR1 = false>
__device__ __forceinline__ void logmean2(float a, float b, float& lm, float& inv_lm) {
    const float s = a + b;
    const float r = __builtin_amdgcn_rcpf(s);
    const float d = a - b;
    const float f = d * r;
    const float u = f * f;
    const bool series = u < 0.2f;
    const float F = 1.0f + u * (1.0f / 3.0f + u * (1.0f / 5.0f + u * (1.0f / 7.0f + u * (1.0f / 9.0f +
                    u * (1.0f / 11.0f + u * (1.0f / 13.0f + u * (1.0f / 15.0f + u * (1.0f / 17.0f +
                    u * (1.0f / 19.0f)))))))));
    lm = 0.5f * s * __builtin_amdgcn_rcpf(F);
    inv_lm = 2.0f * F * r;
    if (__builtin_expect(__any(!series), 0)) {
        const float l = flog(a * __builtin_amdgcn_rcpf(b));
        if (!series) {
            lm = d * __builtin_amdgcn_rcpf(l);
            inv_lm = l * __builtin_amdgcn_rcpf(d);
        }
    }
}

// FP32 log-mean with both forms evaluated and selected (straight-line code):
// the same series (eight terms for f^2 < 0.1, truncation < 1e-9) and log
// form as logmean2, for the GLM-MHD direction pass, where the branch of
// logmean2 raises the register pressure past the register file (1 KB of
// scratch per lane, 2.6x slower).
__device__ __forceinline__ void logmean2_select(float a, float b, float& lm, float& inv_lm) {
    const float r = __builtin_amdgcn_rcpf(a + b);
    const float d = a - b;
    const float f = d * r;
    const float u = f * f;
    const bool series = u < 0.1f;
    const float F = 1.0f + u * (1.0f / 3.0f + u * (1.0f / 5.0f + u * (1.0f / 7.0f + u * (1.0f / 9.0f +
                    u * (1.0f / 11.0f + u * (1.0f / 13.0f + u * (1.0f / 15.0f)))))));
    const float l = flog(a * __builtin_amdgcn_rcpf(b));
    const float num = series ? 0.5f * (a + b) : d;
    const float den = series ? F : l;
    lm = num * __builtin_amdgcn_rcpf(den);
    inv_lm = series ? 2.0f * F * r : l * __builtin_amdgcn_rcpf(d);
}

// Chandrashekar entropy-conserving, kinetic-energy-preserving flux:
// two_point_flux_mod.f90:34-77, rotated to direction D (:79-141).  NEAR1 as
// for logmean2, for the ratios of both densities and both betas.
template <int D, bool NEAR1 = false, typename Real>
__device__ __forceinline__ void ec_flux(const Gas<Real>& g, const NodeState<Real>& a, const NodeState<Real>& b,
                                        Real* f) {
    constexpr int N0 = D, N1 = (D + 1) % 3, N2 = (D + 2) % 3;
    Real dens, inv_dens, lnb, inv_lnb;
    logmean2<NEAR1>(a.u[DENS], b.u[DENS], dens, inv_dens);
    logmean2<NEAR1>(a.beta, b.beta, lnb, inv_lnb);
    const Real velx = Real(0.5) * (a.v[N0] + b.v[N0]);
    const Real vely = Real(0.5) * (a.v[N1] + b.v[N1]);
    const Real velz = Real(0.5) * (a.v[N2] + b.v[N2]);
    const Real pres = fdiv(Real(0.5) * (a.u[DENS] + b.u[DENS]), a.beta + b.beta);
    const Real ssum = (a.v[N0] * a.v[N0] + b.v[N0] * b.v[N0]) + (a.v[N1] * a.v[N1] + b.v[N1] * b.v[N1]) +
                      (a.v[N2] * a.v[N2] + b.v[N2] * b.v[N2]);
    const Real enth = g.half_inv_km1 * inv_lnb + pres * inv_dens + velx * velx + vely * vely + velz * velz -
                      Real(0.25) * ssum;
    const Real fr = dens * velx;
    f[DENS] = fr;
    f[1 + N0] = fr * velx + pres;
    f[1 + N1] = fr * vely;
    f[1 + N2] = fr * velz;
    f[ENER] = fr * enth;
}

// Central two-point flux: equations/euler/two_point/standard/two_point_flux_mod.f90:13-23
template <int D, typename Real>
__device__ __forceinline__ void central_flux(const NodeState<Real>& a, const NodeState<Real>& b, Real* r) {
    Real fu[5], fv[5];
    flux_s<D>(a, fu);
    flux_s<D>(b, fv);
#pragma unroll
    for (int h = 0; h < 5; ++h) r[h] = Real(0.5) * (fu[h] + fv[h]);
}

// isvalid_elemental: equations_mod.f90:337-344
template <typename Real>
__device__ __forceinline__ bool isvalid(const Gas<Real>& g, const Real* u) {
    return (u[DENS] > Real(0)) && (pressure(g, u) > Real(0));
}

}  // namespace nemo
