#!/usr/bin/env python3
"""Writes tests/golden/ref_pointwise.npz: input/output vectors of the
reference's own Fortran leaf functions, evaluated through oracle/_ref (the
reference's modules compiled in place from its source tree by
oracle/ref.mk; nothing of the reference is copied).

Covered functions (reference file:line):
  pressure            equations/euler/equations_mod.f90:131-151
  x/y/zflux           equations/euler/equations_mod.f90:64-127
  x/y/zlambdaMax      equations/euler/equations_mod.f90:384-493
  cons2evec           equations/euler/equations_mod.f90:261-286
  HLLC/HLLE hybrid    equations/euler/riemann/hllc/riemann_mod.f90:7-166
  Rusanov             equations/euler/riemann/rusanov/riemann_mod.f90:7-101
  Chandrashekar       equations/euler/two_point/chandrashekar/two_point_flux_mod.f90:13-141
  central             equations/euler/two_point/standard/two_point_flux_mod.f90:13-51
  LGL nodes/weights   utils/interpol_mod.f90:513-582
  DiffMatrix          utils/interpol_mod.f90:38-80
  Vandermonde (+inv)  utils/interpol_mod.f90:274-311, 584-601; utils/linalg_mod.f90:123-200
  LSRK 5-4 constants  timedisc/rungekutta/low-storage/4-5/timedisc_const_mod.f90:5-29
  SSPRK(5,4) consts   timedisc/rungekutta/ssp/4-5/timedisc_const_mod.f90:5-36
  minmod3             utils/math_mod.f90:77-89

Run: make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def random_states(rng, n, gamma=1.4):
    """Physical states (rho > 0, p > 0) over several orders of magnitude,
    plus near-equal pairs that hit the log-mean series branch."""
    rho = 10.0 ** rng.uniform(-2, 1, n)
    v = rng.uniform(-3, 3, (n, 3))
    p = 10.0 ** rng.uniform(-3, 1, n)
    E = p / (gamma - 1) + 0.5 * rho * (v ** 2).sum(1)
    return np.stack([rho, rho * v[:, 0], rho * v[:, 1], rho * v[:, 2], E], 1)


def main():
    R = O.Ref()
    rng = np.random.default_rng(20261202)
    n = 400
    uL = random_states(rng, n)
    uR = random_states(rng, n)
    # near-equal pairs: series branch of logmean (|a-b| < 1e-4)
    k = n // 4
    uR[:k] = uL[:k] * (1.0 + rng.uniform(-5e-5, 5e-5, (k, 1)))
    # supersonic pairs for the HLLC branches S_L >= 0 / S_R <= 0
    uL[k:k + 20, 1] = uL[k:k + 20, 0] * 8.0
    uR[k:k + 20, 1] = uR[k:k + 20, 0] * 8.0
    uL[k + 20:k + 40, 1] = -uL[k + 20:k + 40, 0] * 8.0
    uR[k + 20:k + 40, 1] = -uR[k + 20:k + 40, 0] * 8.0
    for a in (uL, uR):
        a[k:k + 40, 4] = 0.0
        rho = a[k:k + 40, 0]
        v = a[k:k + 40, 1:4] / rho[:, None]
        a[k:k + 40, 4] = 1.0 / 0.4 + 0.5 * rho * (v ** 2).sum(1)

    out = {"uL": uL, "uR": uR}
    out["pressure"] = np.array([R.pressure(u) for u in uL])
    out["cons2evec"] = np.array([R.cons2evec(u) for u in uL])
    for d in range(3):
        out[f"flux{d}"] = np.array([R.flux(d, u) for u in uL])
        out[f"lmax{d}"] = np.array([R.lambda_max(d, u) for u in uL])
        out[f"hllc{d}"] = np.array([R.hllc(d, a, b) for a, b in zip(uL, uR)])
        out[f"rusanov{d}"] = np.array([R.rusanov(d, a, b) for a, b in zip(uL, uR)])
        out[f"ec{d}"] = np.array([R.ec(d, a, b) for a, b in zip(uL, uR)])
        out[f"central{d}"] = np.array([R.central(d, a, b) for a, b in zip(uL, uR)])
    for N in range(1, 8):
        x, w = R.lgl(N)
        out[f"lgl_x{N}"] = x
        out[f"lgl_w{N}"] = w
        out[f"D{N}"] = R.diff_matrix(N, x)
        V, Vi = R.vandermonde(x)
        out[f"V{N}"] = V
        out[f"Vinv{N}"] = Vi
    a, b, c = R.lsrk_coeffs()
    out["lsrk_a"], out["lsrk_b"], out["lsrk_c"] = a, b, c
    out["ssp"] = R.ssp_coeffs()
    mm = rng.uniform(-1, 1, (200, 3))
    out["minmod3_in"] = mm
    out["minmod3"] = np.array([R.minmod3(*r) for r in mm])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_pointwise.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
