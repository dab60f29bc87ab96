"""The Fortran host driver (nemo_amd/fortran) over the C-ABI.

GPU: the driver runs the reference's loop (calc_timestep, five stages per
step, status check) through the ISO_C_BINDING interfaces for four setups
(Euler isentropic vortex, Euler Kelvin-Helmholtz, GLM-MHD Orszag-Tang, and
Sod's shock tube between slip walls);
its dumped initial state is advanced by the oracle for the same number of
steps and compared with the dumped final state (and the time steps).
CPU: without a device the driver reports the C-ABI error and stops with a
nonzero code (the reference's raise_error/rt%stat convention).
"""
import os
import subprocess

import numpy as np
import pytest

from nemo_amd import mesh as meshmod
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nemo_amd", "nemo_gpu_driver")


def _read_state(path):
    raw = open(path, "rb").read()
    nelem, n_poly, nvars = np.frombuffer(raw[:24], dtype=np.int64)
    u = np.frombuffer(raw[24:], dtype=np.float64).reshape(nelem, nvars, (n_poly + 1) ** 3)
    return np.array(u), int(n_poly)


def _have_exe():
    return os.path.exists(EXE)


@pytest.mark.skipif(not _have_exe(), reason="Fortran driver not built")
def test_driver_reports_errors_without_device(tmp_path):
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"  # no device visible: creation must fail cleanly
    r = subprocess.run([EXE, "vortex", "2", "3", "1", str(tmp_path / "x")], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode != 0
    assert "[ERROR] nemo_gpu_create" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not _have_exe(), reason="Fortran driver not built")
@pytest.mark.parametrize("setup", ["vortex", "kh", "ot", "sod"])
def test_fortran_driver_matches_oracle(tmp_path, setup):
    n, N, steps = 8, 3, 20
    prefix = str(tmp_path / setup)
    print(f"fortran driver {setup} n={n} N={N} steps={steps}", flush=True)
    r = subprocess.run([EXE, setup, str(n), str(N), str(steps), prefix], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    u0, n_poly = _read_state(prefix + ".init.bin")
    u1, _ = _read_state(prefix + ".final.bin")
    dts = np.loadtxt(prefix + ".dt.txt")
    assert n_poly == N and len(dts) == steps
    m = meshmod.uniform_periodic(n)
    if setup == "sod":
        # slip walls across x, built here as the driver builds them
        m = meshmod.with_boundaries(m, [meshmod.BC_SLIP_WALL] * 2 + [None] * 4)
    if setup == "vortex":
        cfg = O.make_cfg(N=N, volume=O.VOL_WEAK, surface=O.SURF_RUSANOV, timedisc=O.TD_SSPRK54)
    elif setup in ("kh", "sod"):
        cfg = O.make_cfg(N=N, volume=O.VOL_EC, surface=O.SURF_HLLC, indicator=O.IND_PP, timedisc=O.TD_LSRK54,
                         repair=1)
        if setup == "sod":
            cfg = O.with_boundaries(cfg, m.bfaces)
    else:
        # GLM-MHD: the oracle's run sets the cleaning speed and damping from
        # each step's CFL reduction, as the device does in calc_timestep
        cfg = O.make_cfg(N=N, eq=O.EQ_MHD, volume=O.VOL_EC, surface=O.SURF_HLLD, indicator=O.IND_PP,
                         timedisc=O.TD_LSRK54, repair=1, gamma=5.0 / 3.0, mu0=1.0)
    assert u0.shape[1] == O.nvars(cfg)
    uref = np.ascontiguousarray(u0)
    rc, href = O.run(cfg, m.nbr, m.delta, uref, steps)
    assert rc == 0
    assert np.abs(dts - href).max() <= 1e-13 * href.max()
    # per variable, relative to its own largest value (psi grows from zero),
    # floored at 1e-3 of the state's largest value (the vortex's z momentum
    # is rounding noise around zero)
    floor = 1e-3 * np.abs(uref).max()
    err = max(float(np.abs(u1[:, h] - uref[:, h]).max() / max(np.abs(uref[:, h]).max(), floor))
              for h in range(u0.shape[1]))
    print(f"  rel Linf after {steps} steps (worst variable) {err:.3e}")
    assert err <= 1e-10
