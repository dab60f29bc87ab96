"""Device volume integrals (nemo_gpu_analyze) against host quadrature sums
and closed forms, on uniform and adaptive meshes, Euler and GLM-MHD."""
import os

import numpy as np
import pytest

from nemo_amd import _lib, amr, analyze, mesh as meshmod, setups, solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _host(mesh, N, u, gamma, mhd=False):
    vol = analyze.quadrature_volumes(mesh, N)
    V = u.shape[1]
    out = {h: float((vol * u[:, h]).sum()) for h in range(V)}
    m2 = (u[:, 1:4] ** 2).sum(1)
    out["ekin"] = float((vol * 0.5 * m2 / u[:, 0]).sum())
    if mhd:
        B2 = (u[:, 5:8] ** 2).sum(1)
        p = (gamma - 1) * (u[:, 4] - 0.5 * m2 / u[:, 0] - 0.5 * B2 - 0.5 * u[:, 8] ** 2)
        out["emag"] = float((vol * 0.5 * B2).sum())
    else:
        p = (gamma - 1) * (u[:, 4] - 0.5 * m2 / u[:, 0])
    out["eint"] = float((vol * p / (gamma - 1)).sum())
    out["rho_min"], out["p_min"] = float(u[:, 0].min()), float(p.min())
    out["v_max"] = float((np.sqrt(m2) / u[:, 0]).max())
    return out


NAMES = ["dens", "momx", "momy", "momz", "ener", "magx", "magy", "magz", "glmp"]


@pytest.mark.parametrize("case", ["kh", "amr", "mhd"])
def test_device_integrals_match_host(case):
    N = 3
    if case == "kh":
        m = meshmod.uniform_periodic(8)
        u = setups.kelvin_helmholtz(m, N, seed=3)
        cfg = solver.make_config(N=N, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP)
        gamma, mhd = 1.4, False
    elif case == "amr":
        m = amr.Forest.uniform(2, maxlevel=3).refine(np.arange(64) % 5 == 0).balance()
        u = setups.kelvin_helmholtz(meshmod.uniform_periodic(1), N, seed=3)[:1].repeat(m.nelem, 0)
        x, _ = O.lgl(N)
        X = m.node_coords(x)
        u[:, 0] = 1.0 + 0.5 * X[:, 0] * X[:, 1] * X[:, 2]    # integral 1 + 1/16
        cfg = solver.make_config(N=N, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP)
        gamma, mhd = 1.4, False
    else:
        m, u = setups.config_c3(n=4, N=N)
        cfg = solver.make_config(N=N, equations=_lib.EQ_GLM_MHD, volume=_lib.VOL_EC, surface=_lib.SURF_HLLD,
                                 indicator=_lib.IND_PP, gamma=5.0 / 3.0)
        gamma, mhd = 5.0 / 3.0, True
    with solver.Solver(cfg) as s:
        s.upload_mesh(m)
        s.set_state(u)
        a = s.analyze()
        a2 = s.analyze()
    assert a == a2  # deterministic
    h = _host(m, N, u, gamma, mhd)
    for k in range(u.shape[1]):
        assert abs(a[NAMES[k]] - h[k]) <= 1e-13 * max(1.0, abs(h[k])), (NAMES[k], a[NAMES[k]], h[k])
    for k in ("ekin", "eint") + (("emag",) if mhd else ()):
        assert abs(a[k] - h[k]) <= 1e-13 * abs(h[k]), k
    for k in ("rho_min", "p_min", "v_max"):
        assert abs(a[k] - h[k]) <= 1e-15 * abs(h[k]), k
    assert abs(a["volume"] - 1.0) < 1e-14
    if case == "amr":
        assert abs(a["dens"] - (1.0 + 0.5 / 8.0)) < 1e-14


def test_integrals_conserved_by_run_and_log(tmp_path):
    N = 3
    m = meshmod.uniform_periodic(8)
    u = setups.kelvin_helmholtz(m, N, seed=3)
    cfg = solver.make_config(N=N, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP,
                             timedisc=_lib.TD_LSRK54, repair=1)
    log = analyze.AnalysisLog(str(tmp_path / "analysis.dat"))
    with solver.Solver(cfg) as s:
        s.upload_mesh(m)
        s.set_state(u)
        a0 = log.record(s, 0.0, 0.0, 0, force=True)
        s.run(20)
        a1 = log.record(s, 1.0, 0.0, 20, force=True)
    for k in ("dens", "momx", "momy", "momz", "ener"):
        assert abs(a1[k] - a0[k]) <= 1e-13 * max(1.0, abs(a0[k])), k
    rows = np.loadtxt(tmp_path / "analysis.dat")
    assert rows.shape == (2, len(analyze.AnalysisLog.COLUMNS))


def test_mainloop_writes_checkpoints_and_analysis(tmp_path):
    """driver_mod's loop with dt_checkpoint: a FLASH file at t = 0, at every
    dump time and at the end, whose leaf data is the device state at that
    time; the analysis log has one line per step."""
    from nemo_amd import flash
    N = 3
    m = meshmod.uniform_periodic(4)
    u = setups.kelvin_helmholtz(m, N, seed=3)
    cfg = solver.make_config(N=N, volume=_lib.VOL_EC, surface=_lib.SURF_HLLC, indicator=_lib.IND_PP,
                             timedisc=_lib.TD_LSRK54, repair=1)
    log = analyze.AnalysisLog(str(tmp_path / "analysis.dat"))
    lines = []
    with solver.Solver(cfg) as s:
        s.upload_mesh(m)
        s.set_state(u)
        rt = s.mainloop(0.01, dt_checkpoint=0.004, checkpoint_dir=str(tmp_path / "checkpoints"), mesh=m,
                        analysis=log, log=lines.append)
        u_end = s.get_state()
    assert abs(rt.simtime - 0.01) < 1e-14 and rt.chkptcount == 4  # t = 0, 0.004, 0.008, 0.01
    files = sorted(os.listdir(tmp_path / "checkpoints"))
    assert files == [f"chkpt_{i:06d}.f5" for i in range(4)]
    assert len(lines) == 4 and lines[0].startswith("[CHECKPOINT]")
    rows = np.loadtxt(tmp_path / "analysis.dat")
    assert rows.shape[0] == rt.timesteps + 1
    import subprocess
    h5dump = "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        out = str(tmp_path / "d.bin")
        subprocess.run([h5dump, "-d", "/dens", "-b", "LE", "-o", out, str(tmp_path / "checkpoints" / files[-1])],
                       check=True, capture_output=True)
        d = np.fromfile(out, dtype="<f8").reshape(-1, 64)
        assert np.array_equal(d[-m.nelem:], u_end[:, 0])
