"""Group-level routines (galhip_getacc, galhip_stpgrp, galhip_massgn): the
lib15 routines getacc(jst), stpgrp(jst) + offgrd + rezgrp(jst) and
massgn(jst) on a caller's group of particles.

  - getacc against the reference's formulas evaluated in numpy at the
    context's precision on the oracle's cells and weights (orc.weight), the
    oracle's field and fixed-force table: grdacc's four-corner sum in corner
    order, the fixed radial force f(r) x / r, the zone's step factor squared
    (SRC/lib15/getacc.f:35-163, grdacc.f:40-66); bit-exact in FP32, 1e-14 in
    FP64;
  - getacc followed by stpgrp on the particles a step moves equals that step
    (galhip_step) particle for particle, zones included (FP32 and FP64,
    bit for bit), over a zone cycle;
  - massgn over the particles, in two groups, equals masset's zone sums and
    ring ranges (integers, exact).
"""
import numpy as np
import pytest

from tests._model import SMALL, gpu_model, oracle_model

pytestmark = pytest.mark.gpu

PRECS = [("fp32", "f"), ("fp64", "d")]


def _say(*a):
    print("[case]", *a, flush=True)


def _mzone(istep, nstep):
    m = 1
    for i, s in enumerate(nstep, start=1):
        if istep % s == 0:
            m = i
    return m


def _getacc_ref(orc, m, x, y, iz, fld, dt):
    """getacc in numpy at precision dt: grid sum in corner order, fixed
    force, zone factor."""
    nr, na, nz = m.nr, m.na, len(m.nstep)
    cell, w, _ = orc.weight(x, y, nr, na, "f" if dt == np.float32 else "d")
    ring, spoke = cell // na, cell % na
    s2 = np.where(spoke + 1 == na, 0, spoke + 1)
    ic = [ring * na + spoke, ring * na + s2, (ring + 1) * na + spoke, (ring + 1) * na + s2]
    mesh = nr * na
    ax = np.zeros(len(x), dt)
    ay = np.zeros(len(x), dt)
    for q in range(4):
        inside = ic[q] < mesh
        idx = np.where(inside, ic[q], 0)
        fx = np.where(inside, fld[0][idx], dt(0)).astype(dt)
        fy = np.where(inside, fld[1][idx], dt(0)).astype(dt)
        ax = (ax + (fx * w[:, q]).astype(dt)).astype(dt)
        ay = (ay + (fy * w[:, q]).astype(dt)).astype(dt)
    x = x.astype(dt)
    y = y.astype(dt)
    rr = np.sqrt((x * x).astype(dt) + (y * y).astype(dt)).astype(dt)
    dr, val, der = m.fixforce
    fr = np.array([orc.fixforce(float(r), dr, val, der)[0] for r in rr], np.float64).astype(dt)
    f = np.where(fr != 0, (fr / np.where(rr > 0, rr, dt(1))).astype(dt), dt(0)).astype(dt)
    halo = fr != 0
    ax = np.where(halo, (ax + (f * x).astype(dt)).astype(dt), ax)
    ay = np.where(halo, (ay + (f * y).astype(dt)).astype(dt), ay)
    off = iz == nz + 1
    ax = np.where(off, dt(0), ax)
    ay = np.where(off, dt(0), ay)
    zz = np.minimum(iz, nz)
    t = np.array(m.nstep, dt)[zz - 1]
    t = (t * t).astype(dt)
    scale = (nz > 1) & (iz > 1)
    ax = np.where(scale, (t * ax).astype(dt), ax)
    ay = np.where(scale, (t * ay).astype(dt), ay)
    return np.stack([ax, ay], axis=1)


@pytest.mark.parametrize("prec,op", PRECS)
def test_getacc_matches_reference_formulas(orc, prec, op):
    _say("group getacc", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    m.masset()
    m.findf()
    dt = np.float32 if op == "f" else np.float64
    fld = m.fld.astype(dt)
    g.set_field(fld[:2])
    out = g.unload()
    x, y, iz = out["x"], out["y"], out["iz"]
    got = g.getacc(x, y, iz)
    ref = _getacc_ref(orc, m, x, y, iz, fld, dt)
    if prec == "fp32":
        assert np.array_equal(got, ref), np.abs(got - ref).max()
    else:
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-14 * scale, np.abs(got - ref).max() / scale
    g.close()


@pytest.mark.parametrize("prec,op", PRECS)
def test_getacc_stpgrp_chain_equals_step(orc, prec, op):
    _say("group getacc + stpgrp vs step", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    m.masset()
    g.masset()
    dt = np.float32 if op == "f" else np.float64
    nz = len(SMALL.nstep)
    for k in range(4):     # a zone cycle: mzone 3, 1, 2, 1 (nstep 1, 2, 4)
        m.findf()
        g.set_field(m.fld.astype(dt)[:2])
        istep = g.get_step()[0]
        mz = _mzone(istep, SMALL.nstep)
        before = g.unload()
        x, y, vx, vy, iz = (before[q] for q in ("x", "y", "vx", "vy", "iz"))
        mv = (iz >= 1) & (iz <= mz)
        if mz == nz:
            mv |= iz == nz + 1
        acc = g.getacc(x[mv], y[mv], iz[mv])
        gx, gy, gvx, gvy, giz = g.stpgrp(x[mv], y[mv], vx[mv], vy[mv], acc, iz[mv])
        g.step()
        m.step(field=m.fld.astype(dt)[:2])
        after = g.unload()
        for q, v in (("x", gx), ("y", gy), ("vx", gvx), ("vy", gvy), ("iz", giz)):
            assert np.array_equal(after[q][mv], v), (k, q)
            assert np.array_equal(after[q][~mv], before[q][~mv]), (k, q)
        _say("step %d (mzone %d): %d particles moved, %d changed list"
             % (istep, mz, int(mv.sum()), int((giz != iz[mv]).sum())))
    g.close()


@pytest.mark.parametrize("prec,op", PRECS)
def test_massgn_groups_equal_masset(orc, prec, op):
    _say("group massgn vs masset", prec)
    m, parts = oracle_model(orc, prec=op)
    g = gpu_model(precision=prec, parts=parts)
    g.masset()
    mfx, _, jr, kr, _ = g.get_mass()
    nz = len(SMALL.nstep)
    g.set_mass(mfx=np.zeros_like(mfx), jrad=np.full(nz, g.nr, np.int32),
               krad=np.zeros(nz, np.int32))
    out = g.unload()
    x, y, iz = out["x"], out["y"], out["iz"]
    h = len(x) // 3
    g.massgn(x[:h], y[:h], iz[:h])
    g.massgn(x[h:], y[h:], iz[h:])
    mfx2, _, jr2, kr2, _ = g.get_mass()
    assert np.array_equal(mfx2, mfx)
    assert np.array_equal(jr2, jr) and np.array_equal(kr2, kr), (jr, jr2, kr, kr2)
    g.close()
