"""GALAXY particle files (.pcs) and the path from one to a start state.

A .pcs file is a Fortran unformatted sequential file with 4-byte record
markers (SRC/lib15/unload.f:104-170, read back by SRC/lib15/loadup.f:52-170):
  record 1: nsp(1:3) (particles per population), ncoor (words per particle,
            ncoor + 1 with individual masses), the block length l (5000),
            the time t, the particle mass pm in model units, the perturber
            flag (a Fortran logical, 4 bytes);
  record 2: the perturber's state, present only when the flag is set;
  then blocks of l particles x ncoor words (x, y, vx, vy in model units for
  2-D), the last block holding the remainder.

to_internal() converts as loadup does, in FP32 (loadup.f:99,168-169):
coordinates times lscale, velocities over gvfac = 1/(lscale ts)
(setrun.f:873), the particle mass pm lscale^3 ts^2.  zones() assigns the
lists (zoneno.f, the off-grid list beyond rgrid), adjvls() scales the
velocities by the zone step factors (adjvls.f).  start_state() then does
what pcs2dmp does before galaxy's first step with the retarded-velocity
leap-frog: load, masset, a full findf, and tmcent (SRC/lib15/tmcent.f).
"""
from __future__ import annotations

import struct

import numpy as np

PCS_BLOCK = 5000


def _records(buf: bytes):
    out, o = [], 0
    while o < len(buf):
        (n,) = struct.unpack("<i", buf[o:o + 4])
        (m,) = struct.unpack("<i", buf[o + 4 + n:o + 8 + n])
        if m != n:
            raise ValueError("bad record markers at byte %d" % o)
        out.append(buf[o + 4:o + 4 + n])
        o += 8 + n
    return out


def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def read_pcs(path) -> dict:
    """The header fields and coords[n, ncoor] (float32, model units)."""
    with open(path, "rb") as f:
        recs = _records(f.read())
    h = recs[0]
    nsp = np.frombuffer(h[0:12], "<i4").copy()
    ncoor, block = (int(v) for v in np.frombuffer(h[12:20], "<i4"))
    t, pm = (float(v) for v in np.frombuffer(h[20:28], "<f4"))
    pertbn = bool(np.frombuffer(h[28:32], "<i4")[0])
    k = 1
    perturber = None
    if pertbn:
        perturber = recs[1]
        k = 2
    data = np.concatenate([np.frombuffer(r, "<f4") for r in recs[k:]]) \
        if len(recs) > k else np.zeros(0, "<f4")
    n = int(nsp.sum())
    if data.size != n * ncoor:
        raise ValueError("%d particles x %d words expected, %d words found"
                         % (n, ncoor, data.size))
    return dict(nsp=nsp, ncoor=ncoor, block=block, t=t, pm=pm, pertbn=pertbn,
                perturber=perturber, coords=data.reshape(n, ncoor).astype(np.float32))


def write_pcs(path, coords, t, pm, nsp=None, block=PCS_BLOCK, perturber=None):
    """Writes coords[n, ncoor] (model units) in the layout read_pcs reads."""
    coords = np.ascontiguousarray(coords, "<f4")
    n, ncoor = coords.shape
    nsp = np.array([n, 0, 0] if nsp is None else nsp, "<i4")
    if int(nsp.sum()) != n:
        raise ValueError("population counts do not add up to the particles")
    head = (nsp.tobytes() + np.array([ncoor, block], "<i4").tobytes() +
            np.array([t, pm], "<f4").tobytes() +
            np.array([1 if perturber is not None else 0], "<i4").tobytes())
    with open(path, "wb") as f:
        f.write(_record(head))
        if perturber is not None:
            f.write(_record(bytes(perturber)))
        for s in range(0, n, block):
            f.write(_record(coords[s:s + block].tobytes()))


def to_internal(coords, lscale, ts):
    """Grid-unit positions and per-step velocities, FP32 as loadup forms
    them: x * lscale, v / gvfac with gvfac = 1 / (lscale * ts)."""
    f32 = np.float32
    ls, tsf = f32(lscale), f32(ts)
    gvfac = f32(1.0) / (ls * tsf)
    c = np.asarray(coords, f32)
    x, y = c[:, 0] * ls, c[:, 1] * ls
    vx, vy = c[:, 2] / gvfac, c[:, 3] / gvfac
    return x, y, vx, vy


def internal_mass(pm, lscale, ts):
    """pmass = pm * lscale**3 * ts**2 in FP32 (loadup.f:99)."""
    f32 = np.float32
    ls, tsf = f32(lscale), f32(ts)
    return float(f32(pm) * (ls * ls * ls) * (tsf * tsf))


def zones(x, y, zbr, rgrid):
    """List of each particle: the last zone whose inner radius zbr (grid
    units, zbr[0] = 0) the FP32 radius reaches (zoneno.f:40-55), or the
    off-grid list len(zbr) + 1 beyond rgrid (offgrd.f:55-59)."""
    f32 = np.float32
    x = np.asarray(x, f32)
    y = np.asarray(y, f32)
    r = np.sqrt(x * x + y * y)
    iz = np.ones(len(r), np.int32)
    for i, zb in enumerate(zbr[1:], start=2):
        iz[r >= f32(zb)] = i
    iz[r > f32(rgrid)] = len(zbr) + 1
    return iz


def adjvls(vx, vy, iz, nstep):
    """Velocities times the step factor of their zone (adjvls.f); the
    off-grid list is left alone."""
    f32 = np.float32
    fac = np.ones(len(iz), f32)
    for z, s in enumerate(nstep, start=1):
        fac[iz == z] = f32(s)
    return np.asarray(vx, f32) * fac, np.asarray(vy, f32) * fac


def start_state(grid, coords, lscale, ts, zbr, nstep):
    """pcs2dmp's path on a P2DGrid: load the converted particles, assign
    their mass, solve for the full field, and time-centre the velocities.
    Returns the zones assigned."""
    x, y, vx, vy = to_internal(coords, lscale, ts)
    iz = zones(x, y, zbr, grid.rgrid)
    vx, vy = adjvls(vx, vy, iz, nstep)
    grid.restrt(x, y, vx, vy, iz)
    grid.masset()
    grid.findf(False)
    grid.tmcent(True)
    return iz


def end_state(grid):
    """dmp2pcs's path before the particles are written (dmp2pcs.f:89-100):
    assign every particle's mass, solve for the full field with the
    potential, and undo the time-centring as unload does (unload.f:45), so
    the velocities are synchronous with the positions again.  Returns the
    particles in load order (P2DGrid.unload)."""
    grid.masset()
    grid.findf(True)
    grid.tmcent(False)
    return grid.unload()
