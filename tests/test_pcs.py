"""The .pcs format and the path from a .pcs file to the time-centred start
state, against the reference's own files: EXAMPLES/p2d/run100.pcs0 (copied
verbatim as tests/golden/run100.pcs0) and the restart dump the reference's
pcs2dmp made from it, run100.dmp0 (tests/golden/run100_dmp0.npz).

  - read_pcs / write_pcs round-trip the file byte for byte;
  - the FP32 conversion to grid units (loadup.f:99,168-169) gives the dump's
    positions and particle mass bit for bit, and the zone lists match;
  - the oracle's masset + findf + tmcent gives the dump's velocities to
    within one FP32 ulp of the largest velocity (the field behind the half-
    step correction differs from the reference's FP32 FFT by rounding only);
  - on the GPU, galhip_tmcent with the oracle's field is bit-identical to
    the oracle's tmcent, and the whole GPU path meets the dump as above.
"""
import os

import numpy as np
import pytest

from galaxy_amd import pcs
from tests._model import kt_setup

HERE = os.path.dirname(os.path.abspath(__file__))
PCS0 = os.path.join(HERE, "golden", "run100.pcs0")


def _dump():
    return np.load(os.path.join(HERE, "golden", "run100_dmp0.npz"))


def test_pcs_header_and_roundtrip(tmp_path):
    p = pcs.read_pcs(PCS0)
    assert list(p["nsp"]) == [15000, 0, 0]
    assert p["ncoor"] == 4 and p["block"] == 5000 and not p["pertbn"]
    assert p["t"] == 0.0
    assert abs(p["pm"] * 15000 / 0.857301 - 1) < 1e-6   # run100.lis: active mass 0.857301
    out = tmp_path / "copy.pcs"
    pcs.write_pcs(out, p["coords"], p["t"], p["pm"], nsp=p["nsp"], block=p["block"])
    assert out.read_bytes() == open(PCS0, "rb").read()


def test_pcs_partial_block(tmp_path):
    c = np.arange(7 * 5, dtype=np.float32).reshape(7, 5)
    f = tmp_path / "x.pcs"
    pcs.write_pcs(f, c, 1.5, 0.25, block=3)
    q = pcs.read_pcs(f)
    assert np.array_equal(q["coords"], c) and q["ncoor"] == 5 and q["t"] == 1.5


def test_conversion_matches_dump():
    d = _dump()
    rgrid, lscale, ts, pmass, _ = d["reals"]
    p = pcs.read_pcs(PCS0)
    x, y, vx, vy = pcs.to_internal(p["coords"], lscale, ts)
    assert np.array_equal(x, d["x"]) and np.array_equal(y, d["y"])
    assert np.float32(pcs.internal_mass(p["pm"], lscale, ts)) == np.float32(pmass)
    kw, _, _, _ = kt_setup()
    iz = pcs.zones(x, y, kw["zbr"], rgrid)
    assert np.array_equal(iz, d["iz"])


def _oracle_start(orc):
    d = _dump()
    _, lscale, ts, _, _ = d["reals"]
    kw, ff, _, _ = kt_setup()
    p = pcs.read_pcs(PCS0)
    x, y, vx, vy = pcs.to_internal(p["coords"], lscale, ts)
    iz = pcs.zones(x, y, kw["zbr"], d["reals"][0])
    vx, vy = pcs.adjvls(vx, vy, iz, kw["nstep"])
    m = orc.P2D(fixforce=ff, prec="f", use_fx=True, **kw)
    m.load(x, y, vx, vy, iz)
    m.masset()
    m.findf(False)
    return m, d


def _close_to_dump(vx, vy, d):
    ulp = np.spacing(np.float32(np.abs(d["vx"]).max()))
    assert np.abs(vx - d["vx"]).max() <= ulp
    assert np.abs(vy - d["vy"]).max() <= ulp
    assert (vx == d["vx"]).mean() > 0.95 and (vy == d["vy"]).mean() > 0.95


def test_oracle_tmcent_matches_dump(orc):
    m, d = _oracle_start(orc)
    m.tmcent(True)
    _close_to_dump(m.vx, m.vy, d)


@pytest.mark.gpu
def test_gpu_tmcent_bit_exact_and_dump(orc):
    from galaxy_amd.lib15 import P2DGrid
    print("[pcs] tmcent", flush=True)
    m, d = _oracle_start(orc)
    kw, ff, _, _ = kt_setup()
    g = P2DGrid(precision="fp32", **kw)
    g.greenm()
    g.set_fixforce(*ff)
    g.restrt(m.x, m.y, m.vx, m.vy, m.iz)
    g.masset()
    g.findf(False)
    # the oracle's field in: the half-step correction is then bit-exact
    fld = np.zeros((3, kw["nr"] * kw["na"]), np.float32)
    fld[:2] = m.fld[:2]
    g.set_field(fld)
    g.tmcent(True)
    m.tmcent(True)
    out = g.unload()
    assert np.array_equal(out["vx"], m.vx) and np.array_equal(out["vy"], m.vy)
    with pytest.raises(Exception):
        g.tmcent(True)                  # the reference's logic-error case
    g.close()
    # the GPU's own path from the file
    g = P2DGrid(precision="fp32", **kw)
    g.greenm()
    g.set_fixforce(*ff)
    _, lscale, ts, _, _ = d["reals"]
    pcs.start_state(g, pcs.read_pcs(PCS0)["coords"], lscale, ts, kw["zbr"], kw["nstep"])
    out = g.unload()
    _close_to_dump(out["vx"], out["vy"], d)
    g.close()
