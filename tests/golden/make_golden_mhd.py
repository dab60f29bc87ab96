#!/usr/bin/env python3
"""Writes tests/golden/ref_mhd_pointwise.npz: input/output vectors of the
reference's own GLM-MHD Fortran functions, evaluated through oracle/_ref
library C (the reference's modules compiled in place from its source tree by
oracle/ref.mk with PP_N_VARS=9 and the Orszag-Tang setup_config_mod:
kappa = 5/3, mu0 = 1; nothing of the reference is copied).

Covered functions (reference file:line):
  pressure              equations/mhd/equations_mod.f90:223-245
  x/y/zflux (GLM)       equations/mhd/equations_mod.f90:64-220
  Fastest_Signal_Speed  equations/mhd/equations_mod.f90:693-714
  x/y/zlambdaMax        equations/mhd/equations_mod.f90:522-690
  isvalid               equations/mhd/equations_mod.f90:467-485
  Rusanov               equations/mhd/riemann/rusanov/riemann_mod.f90:7-101
  Powell/GLM NC terms   equations/mhd/non_conservative_flux_mod.f90:10-205

The EC-MHD two-point flux (equations/mhd/two_point/ec) and HLLD
(equations/mhd/riemann/hlld) do not compile as shipped, so they have no
vectors here; tests/test_oracle_mhd.py pins them by their defining
properties instead.

Run: make -C oracle ref && python tests/golden/make_golden_mhd.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

GAMMA = 5.0 / 3.0
GLM_CH = 0.8125


def random_mhd_states(rng, n):
    """Physical GLM-MHD states with plasma beta from 1e-2 to 1e2."""
    rho = 10.0 ** rng.uniform(-1.5, 1, n)
    v = rng.uniform(-2, 2, (n, 3))
    B = rng.uniform(-2, 2, (n, 3))
    psi = rng.uniform(-0.3, 0.3, n)
    p = 10.0 ** rng.uniform(-2, 1, n)
    E = p / (GAMMA - 1) + 0.5 * rho * (v ** 2).sum(1) + 0.5 * ((B ** 2).sum(1) + psi ** 2)
    return np.column_stack([rho, rho[:, None] * v, E, B, psi])


def main():
    R = O.RefMHD()
    R.set_glm(GLM_CH)
    rng = np.random.default_rng(20261203)
    n = 300
    uL, uR, uI = random_mhd_states(rng, n), random_mhd_states(rng, n), random_mhd_states(rng, n)
    # B_x = 0 and B = 0 states (degenerate fast speeds)
    uL[:10, 5] = 0.0
    uL[10:20, 5:8] = 0.0
    for a in (uL,):
        rho = a[:20, 0]
        v = a[:20, 1:4] / rho[:, None]
        a[:20, 4] = 1.0 / (GAMMA - 1) + 0.5 * rho * (v ** 2).sum(1) + 0.5 * ((a[:20, 5:8] ** 2).sum(1) + a[:20, 8] ** 2)
    out = {"uL": uL, "uR": uR, "uI": uI, "gamma": np.array(GAMMA), "glm_ch": np.array(GLM_CH)}
    out["pressure"] = np.array([R.pressure(u) for u in uL])
    out["fastspeed"] = np.array([R.fastspeed(u) for u in uL])
    out["isvalid"] = np.array([R.isvalid(u) for u in uL])
    for d in range(3):
        out[f"flux{d}"] = np.array([R.flux(d, u) for u in uL])
        out[f"lmax{d}"] = np.array([R.lambda_max(d, u) for u in uL])
        out[f"rusanov{d}"] = np.array([R.rusanov(d, a, b) for a, b in zip(uL, uR)])
        out[f"nc{d}"] = np.array([R.nc(d, i, a, b) for i, a, b in zip(uI, uL, uR)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_mhd_pointwise.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
