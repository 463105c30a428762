"""The native driver (galaxy_amd/galhip_galaxy, csrc/galhip_galaxy.cpp):
pcs2dmp's start, galaxy's loop and dmp2pcs's end through the C ABI from
C++, against the same calls made from Python (galaxy_amd.pcs.start_state +
P2DGrid.run + pcs.end_state) on the same .pcs file: the end states must be
bit-identical.  With no steps, the written velocities are the input's to
rounding (time-centring undone, SRC/lib15/unload.f:45)."""
import os
import subprocess

import numpy as np
import pytest

from galaxy_amd import configs, disc, pcs
from galaxy_amd.lib15 import P2DGrid
from tests._model import SMALL

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "galaxy_amd", "galhip_galaxy")


def _run_driver(tmp_path, steps):
    w = SMALL
    n = 20000
    R, phi, vR, vphi = disc.exponential_disc(n, 21, Q=w.Q, md=w.md)
    coords = np.stack([R * np.cos(phi), R * np.sin(phi),
                       vR * np.cos(phi) - vphi * np.sin(phi),
                       vR * np.sin(phi) + vphi * np.cos(phi)], axis=1).astype(np.float32)
    pm = np.float32(w.md / n)
    src, dst = str(tmp_path / "in.pcs"), str(tmp_path / "out.pcs")
    pcs.write_pcs(src, coords, 0.0, pm)
    kw = configs.grid_kwargs(w, n)
    ls, ts = w.lscale(), w.dt
    g = P2DGrid(precision="fp32", **dict(kw, pmass=pcs.internal_mass(pm, ls, ts)))
    dr, val, der = configs.fixforce(w, g.rgrid)
    halo = str(tmp_path / "halo.bin")
    with open(halo, "wb") as f:
        f.write(np.int32(len(val)).tobytes() + np.float64(dr).tobytes()
                + np.asarray(val, np.float64).tobytes() + np.asarray(der, np.float64).tobytes())
    args = [DRIVER, src, dst, "nr=%d" % kw["nr"], "na=%d" % kw["na"], "hash=%d" % kw["hash_"],
            "soft=%r" % kw["softl"], "lscale=%r" % ls, "ts=%r" % ts, "steps=%d" % steps,
            "nstep=" + ",".join(str(s) for s in w.nstep),
            "zbr=" + ",".join(repr(z) for z in kw["zbr"]), "sort=3", "halo=" + halo]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = pcs.read_pcs(dst)
    # the same calls from Python
    g.greenm()
    g.set_fixforce(dr, val, der)
    g.set_sort_interval(3)
    pcs.start_state(g, coords, ls, ts, kw["zbr"], w.nstep)
    g.run(steps)
    u = pcs.end_state(g)
    g.close()
    f32 = np.float32
    lsf, tsf = f32(ls), f32(ts)
    gvfac = f32(1.0) / (lsf * tsf)
    fac = np.ones(n, f32)
    for z, s in enumerate(w.nstep, start=1):
        fac[u["iz"] == z] = f32(s)
    ref = np.stack([u["x"] / lsf, u["y"] / lsf, u["vx"] / fac * gvfac, u["vy"] / fac * gvfac],
                   axis=1).astype(np.float32)
    assert out["coords"].shape == ref.shape
    assert np.array_equal(out["coords"], ref)
    assert out["t"] == np.float32(steps * ts)
    assert out["pm"] == pm and list(out["nsp"]) == [n, 0, 0]
    return coords, out


def test_driver_matches_python_path(tmp_path):
    print("[case] native driver, 12 steps", flush=True)
    _run_driver(tmp_path, 12)


def test_driver_no_steps_returns_input(tmp_path):
    """pcs in, no steps, pcs out: tmcent(.true.) then tmcent(.false.) with
    the same field, so positions come back to one unit-conversion rounding
and velocities come back to
    within a few FP32 roundings of the loaded ones (the half-step kick is
    added and subtracted, each rounded, and the unit conversions round)."""
    print("[case] native driver, no steps", flush=True)
    coords, out = _run_driver(tmp_path, 0)
    c = out["coords"]
    dx = np.abs(c[:, :2] - coords[:, :2]) / np.abs(coords[:, :2]).max()
    assert dx.max() <= 2.5e-7, dx.max()   # x lscale / lscale in FP32
    dv = np.abs(c[:, 2:] - coords[:, 2:]).max()
    assert dv <= 4e-7 * np.abs(coords[:, 2:]).max(), dv
