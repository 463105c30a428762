"""Pins the oracle's weight/massgn/zoncmb/scaled against the reference's
own restart dumps (EXAMPLES/p2d/run100.dmp{0,10,80}, converted to
tests/golden/run100_*.npz by tests/golden/make_run100_fixtures.py).

The dumps come from an unknown platform and compiler, so they pin integer
results exactly (the set of cells that received mass) and floating point
masses to FP32 accumulation-order rounding.
run100.dat: 93 x 128 grid, HASH 2, sectors 2, 2 zones (nstep 1, 2).
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(tag):
    d = np.load(os.path.join(GOLD, "run100_%s.npz" % tag))
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("tag", ["dmp0", "dmp10", "dmp80"])
def test_dump_mass_assignment(orc, tag):
    d = _load(tag)
    nr, na, ng, istep, nbod = (int(v) for v in d["scalars"])
    rgrid, lscale, ts, pmass, softl = d["reals"]
    nsect, nz = 2, 2
    a = orc.alpha(na, "f")
    assert abs(orc.grofu(np.float32(nr - 1), a, "f") - rgrid) == 0.0
    mesh = nr * na
    iz = d["iz"].astype(np.int32)
    grd = np.zeros(nz * mesh, np.float32)
    mfx = np.zeros(nz * mesh, np.int64)
    jr = np.full(nz, nr, np.int32)
    kr = np.zeros(nz, np.int32)
    orc._lib.orc_massgn_f(nbod, d["x"], d["y"], iz, nr, na, a, nsect, nz,
                          grd, mfx, 40, jr, kr)
    m1, m2 = grd[:mesh], grd[mesh:]
    q1 = mfx[:mesh] * 2.0**-40
    q2 = mfx[mesh:] * 2.0**-40
    # zone 2 raw masses: the touched-cell set is exact, values to rounding
    assert np.array_equal(d["grdmss2"] != 0, m2 != 0)
    assert np.abs(d["grdmss2"] - m2).max() < 1e-4
    assert np.abs(d["grdmss2"] - q2).max() < 1e-4
    # zone 1 array after zoncmb (new fraction 1, old skipped) and scaled
    assert d["lstep"][1] == istep
    comb = (m1 + m2) * np.float32(pmass)
    assert np.array_equal(d["grdmss1"] != 0, comb != 0)
    assert np.abs(d["grdmss1"] - comb).max() < 4e-5 * pmass * 10
    assert np.abs(d["grdmss1"] - (q1 + q2) * pmass).max() < 4e-5 * pmass * 10
    # mass conservation: every particle is on the grid (FP32 weights sum
    # to one within a few ulp each)
    assert abs(mfx.sum() * 2.0**-40 - nbod) < 4.8e-7 * nbod
    # all mass folded into the first sector (weight.f:80-81)
    cells = np.nonzero(m1 + m2)[0]
    assert (cells % na).max() <= na // nsect


def test_dump_zone_arrays_consistent(orc):
    """dmp10 was written at istep 500 with lstep(2,2) = 500: its zone-2
    array equals the recomputed zone-2 masses of the dumped positions, and
    the zoncmb fractions (1 for the new array, none for the old) reproduce
    the scaled zone-1 array through the oracle's own bookkeeping."""
    d = _load("dmp10")
    nr, na, ng, istep, nbod = (int(v) for v in d["scalars"])
    rgrid, lscale, ts, pmass, softl = d["reals"]
    m = orc.P2D(nr, na, ng - 1, float(softl), nsect=2, skip=(0,), nzones=2,
                nstep=(1, 2), zbr=(0.0, 30.0), pmass=float(pmass), prec="f",
                use_fx=False)
    m.load(d["x"], d["y"], d["vx"], d["vy"], iz=d["iz"])
    m.istep = istep
    m.lstep[:, 1] = d["lstep"]
    m.mzone = 2
    m._masscl(False)
    m.lstep[:, 1] = d["lstep"]
    fn = orc._lib.orc_massgn_f
    fn(nbod, m.x, m.y, m.iz, nr, na, m.alpha, 2, 2, m.grdmss.reshape(-1),
       m.mfx.reshape(-1), m.qbits, m.jrad, m.krad)
    m._zoncmb_scaled()
    assert np.array_equal(m.rho != 0, d["grdmss1"] != 0)
    assert np.abs(m.rho - d["grdmss1"]).max() < 4e-4 * pmass
