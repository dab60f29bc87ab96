// Host-side helpers shared by the C-ABI translation units (gfx950 build):
// operator construction (own implementations of the reference's algorithms),
// kernel argument tables and the error convention.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/nemo_gpu.h"
#include "dg_stage.hpp"

namespace nemo_host {

// Last error message of the calling thread (nemo_gpu_error_message).
std::string& last_error();

inline int fail(int code, const std::string& msg) {
    last_error() = msg;
    return code;
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return ::nemo_host::fail(NEMO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// Operators on the host (own implementation of the reference's algorithms)
// ---------------------------------------------------------------------------

constexpr int kMaxM = 8;

// Legendre recursion used by the LGL Newton iteration
// (qAndLEvaluation, utils/interpol_mod.f90:468-509).
inline void q_and_L(int N, double x, double& q, double& dq, double& LN) {
    double Lm2 = 1.0, Lm1 = x, dLm2 = 0.0, dLm1 = 1.0, L = 0.0, dL = 0.0;
    for (int k = 2; k <= N; ++k) {
        const double rk = double(k);
        L = (2.0 * rk - 1.0) / rk * x * Lm1 - (rk - 1.0) / rk * Lm2;
        dL = dLm2 + (2.0 * rk - 1.0) * Lm1;
        Lm2 = Lm1;
        Lm1 = L;
        dLm2 = dLm1;
        dLm1 = dL;
    }
    const double rk = double(N + 1);
    const double Lk = (2.0 * rk - 1.0) / rk * x * L - (rk - 1.0) / rk * Lm2;
    const double dLk = dLm2 + (2.0 * rk - 1.0) * Lm1;
    q = Lk - Lm2;
    dq = dLk - dLm2;
    LN = L;
}

// LGL nodes and weights: Newton on q = L_{N+1} - L_{N-1} from Parter's
// asymptotic guess (utils/interpol_mod.f90:513-582).
inline void lgl(int N, double* x, double* w) {
    const double pi = 4.0 * std::atan(1.0);
    double q, dq, LN = 0.0;
    if (N == 1) {
        x[0] = -1.0; x[1] = 1.0; w[0] = w[1] = 1.0;
        return;
    }
    x[0] = -1.0;
    x[N] = 1.0;
    w[0] = w[N] = 2.0 / (double(N) * (double(N) + 1));
    for (int j = 1; j <= (N + 1) / 2 - 1; ++j) {
        const double rj = double(j), rN = double(N);
        x[j] = -std::cos(((rj + 0.25) * pi) / rN - (3.0 / (8.0 * rN * pi)) * (1.0 / (rj + 0.25)));
        for (int it = 0; it <= 4; ++it) {
            q_and_L(N, x[j], q, dq, LN);
            const double d = -q / dq;
            x[j] += d;
            if (std::fabs(d) <= 4.0e-16 * std::fabs(x[j])) break;
        }
        q_and_L(N, x[j], q, dq, LN);
        x[N - j] = -x[j];
        w[j] = w[N - j] = 2.0 / (double(N) * (double(N) + 1.0) * (LN * LN));
    }
    if (N % 2 == 0) {
        q_and_L(N, 0.0, q, dq, LN);
        x[N / 2] = 0.0;
        w[N / 2] = 2.0 / (double(N) * (double(N) + 1.0) * (LN * LN));
    }
}

// Barycentric differentiation matrix (utils/interpol_mod.f90:38-80).
inline void diff_matrix(int N, const double* x, double* D) {
    const int M = N + 1;
    double wb[kMaxM];
    for (int i = 0; i < M; ++i) wb[i] = 1.0;
    for (int i = 1; i < M; ++i)
        for (int j = 0; j < i; ++j) {
            wb[j] *= (x[j] - x[i]);
            wb[i] *= (x[i] - x[j]);
        }
    for (int i = 0; i < M; ++i) wb[i] = 1.0 / wb[i];
    for (int i = 0; i < M * M; ++i) D[i] = 0.0;
    for (int l = 0; l < M; ++l)
        for (int g = 0; g < M; ++g)
            if (l != g) {
                D[g * M + l] = wb[l] / (wb[g] * (x[g] - x[l]));
                D[g * M + g] -= D[g * M + l];
            }
}

// Normalised Legendre polynomial (utils/interpol_mod.f90:274-311).
inline double legendre_hat(int n, double x) {
    double L;
    if (n == 0) {
        L = 1.0;
    } else if (n == 1) {
        L = x;
    } else {
        double Lm2 = 1.0, Lm1 = x;
        L = 0.0;
        for (int i = 2; i <= n; ++i) {
            L = (double(2 * i - 1) * x * Lm1 - double(i - 1) * Lm2) / double(i);
            Lm2 = Lm1;
            Lm1 = L;
        }
    }
    return L * std::sqrt(double(n) + 0.5);
}

// Inverse by Doolittle LU without pivoting (utils/linalg_mod.f90:123-200).
inline void invert(int n, const double* A, double* C) {
    double a[kMaxM * kMaxM], L[kMaxM * kMaxM] = {0}, U[kMaxM * kMaxM] = {0};
    std::memcpy(a, A, sizeof(double) * n * n);
    for (int k = 0; k < n - 1; ++k)
        for (int i = k + 1; i < n; ++i) {
            const double coeff = a[i * n + k] / a[k * n + k];
            L[i * n + k] = coeff;
            for (int j = k + 1; j < n; ++j) a[i * n + j] -= coeff * a[k * n + j];
        }
    for (int i = 0; i < n; ++i) L[i * n + i] = 1.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) U[i * n + j] = a[i * n + j];
    for (int k = 0; k < n; ++k) {
        double d[kMaxM], x[kMaxM];
        for (int i = 0; i < n; ++i) {
            d[i] = (i == k) ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) d[i] -= L[i * n + j] * d[j];
        }
        for (int i = n - 1; i >= 0; --i) {
            x[i] = d[i];
            for (int j = n - 1; j > i; --j) x[i] -= U[i * n + j] * x[j];
            x[i] /= U[i * n + i];
        }
        for (int i = 0; i < n; ++i) C[i * n + k] = x[i];
    }
}

struct HostOps {
    int N, M;
    double x[kMaxM], w[kMaxM], D[kMaxM * kMaxM], Vinv[kMaxM * kMaxM];
};

inline void build_host_ops(int N, HostOps& op) {
    op.N = N;
    op.M = N + 1;
    const int M = op.M;
    lgl(N, op.x, op.w);
    diff_matrix(N, op.x, op.D);
    double V[kMaxM * kMaxM];
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) V[i * M + j] = legendre_hat(j, op.x[i]);
    invert(M, V, op.Vinv);
}

// 2:1 operators of the non-conforming faces and of refinement/coarsening,
// per direction and half c (0 = lower, 1 = upper):
//   P[c][i][p] = l_p((x_i + 2c - 1) / 2)   interpolation of the parent's
//                nodal values to the nodes of half c (the reference's
//                rfaceMat / refineMat, kernel/mcfv/3d/kernel_share_mod.f90:36-90);
//   R[c] = (1/2) Mass^-1 P[c]^T Mass       the L2 projection of a polynomial on
//                half c back to the parent (cfaceMat / coarseMat, :49, which is
//                0.5 rfaceMat^T for the reference's equal-weight subcells).
// Mass is the exact mass matrix of the Lagrange basis, Vinv^T Vinv in the
// orthonormal Legendre basis.  R[0] P[0] + R[1] P[1] = I (coarsening undoes
// refinement) and sum_p w_p R[c][p][i] = w_i / 2 (conservative).
inline void mortar_ops(const HostOps& h, double* P /* [2][M][M] */, double* R /* [2][M][M] */) {
    const int M = h.M;
    for (int c = 0; c < 2; ++c)
        for (int i = 0; i < M; ++i) {
            const double xs = 0.5 * (h.x[i] + (c ? 1.0 : -1.0));
            for (int p = 0; p < M; ++p) {
                double l = 1.0;
                for (int q = 0; q < M; ++q)
                    if (q != p) l *= (xs - h.x[q]) / (h.x[p] - h.x[q]);
                P[(c * M + i) * M + p] = l;
            }
        }
    double mass[kMaxM * kMaxM], minv[kMaxM * kMaxM];
    for (int p = 0; p < M; ++p)
        for (int q = 0; q < M; ++q) {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < M; ++j) {
                a += h.Vinv[j * M + p] * h.Vinv[j * M + q];
                b += legendre_hat(j, h.x[p]) * legendre_hat(j, h.x[q]);
            }
            mass[p * M + q] = a;
            minv[p * M + q] = b;
        }
    for (int c = 0; c < 2; ++c) {
        double t[kMaxM * kMaxM];  // P[c]^T Mass
        for (int a = 0; a < M; ++a)
            for (int i = 0; i < M; ++i) {
                double s = 0.0;
                for (int b = 0; b < M; ++b) s += P[(c * M + b) * M + a] * mass[b * M + i];
                t[a * M + i] = s;
            }
        for (int p = 0; p < M; ++p)
            for (int i = 0; i < M; ++i) {
                double s = 0.0;
                for (int a = 0; a < M; ++a) s += minv[p * M + a] * t[a * M + i];
                R[(c * M + p) * M + i] = 0.5 * s;
            }
    }
}

template <int M, typename Real>
nemo::OpTable<M, Real> make_table(const HostOps& h) {
    nemo::OpTable<M, Real> t;
    for (int i = 0; i < M; ++i) {
        t.w[i] = Real(h.w[i]);
        t.inv_w[i] = Real(1.0 / h.w[i]);
        // weak-form surface lifting on LGL nodes: 2/N_NODES/w_i * l_i(+-1)
        // (kernel/dgsem/weakform/2d/kernel_operators_mod.f90:656-687)
        t.lagrM[i] = Real(i == 0 ? 2.0 / double(M) / h.w[i] : 0.0);
        t.lagrP[i] = Real(i == M - 1 ? 2.0 / double(M) / h.w[i] : 0.0);
        for (int j = 0; j < M; ++j) {
            t.S[i][j] = Real(i == j ? 0.0 : 2.0 * h.D[i * M + j]);
            // weak-form volume matrix: -(-2 w_j/w_i D_ji)/N_NODES
            // (kernel_operators_mod.f90:94-126)
            const double temp = -2.0 * h.w[j] / h.w[i] * h.D[j * M + i];
            t.dgDiff[i][j] = Real(-temp / double(M));
            t.Vinv[i][j] = Real(h.Vinv[i * M + j]);
        }
    }
    return t;
}

inline nemo::RKCoeffs make_rk() {
    nemo::RKCoeffs rk;
    // low-storage/4-5/timedisc_const_mod.f90:5-21 (a_rk, c_rk)
    const double a[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                         -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
    const double c[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};
    // ssp/4-5/timedisc_const_mod.f90:5-32
    const double s[23] = {1.00000000000000, 1.00000000000000, 0.00000000000000, 0.39175222700392,
                          0.39175222700392, 0.44437049406734, 0.55562950593266, 0.36841059262959,
                          0.58607968896780, 0.62010185138540, 0.37989814861460, 0.25189177424738,
                          0.47454236302687, 0.17807995410773, 0.82192004589227, 0.54497475021237,
                          0.93501063100924, 0.00683325884039, 0.34833675773694, 0.22600748319395,
                          0.51723167208978, 0.12759831133288, 0.08460416338212};
    for (int i = 0; i < 5; ++i) {
        rk.lsrk_a[i] = a[i];
        rk.lsrk_c[i] = c[i];
    }
    for (int i = 0; i < 23; ++i) rk.ssp[i] = s[i];
    for (int i = 0; i < 8; ++i) rk.w64[i] = 0.0;
    return rk;
}

// Elements per workgroup: one wavefront of lines (M*M lines per element).
template <int M>
struct Tile;
template <>
struct Tile<3> { static constexpr int BE = 7; };
template <>
struct Tile<4> { static constexpr int BE = 4; };
template <>
struct Tile<5> { static constexpr int BE = 2; };

template <typename Real>
nemo::Gas<Real> make_gas(double k) {
    nemo::Gas<Real> g;
    g.kappa = Real(k);
    g.kappa_m1 = Real(k - 1.0);
    g.hllc_beta = Real(std::sqrt(0.5 * (k - 1) / k));
    g.half_inv_km1 = Real(0.5 / (k - 1.0));
    return g;
}

template <typename Real>
nemo::MhdGas<Real> make_mhd(double k, double mu0) {
    nemo::MhdGas<Real> g;
    g.kappa = Real(k);
    g.kappa_m1 = Real(k - 1.0);
    g.kappa_m2 = Real(k - 2.0);
    g.mu0 = Real(mu0);
    g.smu0 = Real(1.0 / mu0);
    g.half_inv_km1 = Real(0.5 / (k - 1.0));
    g.ch = Real(0);
    return g;
}

inline nemo::Phys64 make_phys64(const nemo_gpu_config& c) {
    nemo::Phys64 p;
    p.gas = make_gas<double>(c.gamma);
    p.mhd = make_mhd<double>(c.gamma, c.mu0);
    return p;
}

}  // namespace nemo_host
