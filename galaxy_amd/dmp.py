"""GALAXY restart dumps (.dmp) for 2-D polar-grid runs.

A .dmp file is a Fortran unformatted sequential file with 4-byte record
markers, written by dumpp (SRC/lib15/dumpp.f:57-198) and read by restrt
(SRC/lib15/restrt.f).  For a p2d run without supplementary forces,
perturber or centring records the records are:
  1-3    the run header (hedrec), kept here as raw bytes;
  4      irun, istep, noff, ncen(mcmp), ... (kept raw except istep, noff);
  5      numprocs, myid, nbod, nwpp, then the head of every particle list
         (nzones zone lists and the off-grid list; -1 = empty), as word
         offsets into the particle array;
  6      the random-generator state (rvsave), raw;
  7      grdmss(:, 1);
  8, 9   per zone z = 2..nzones: lstep(1, z) + grdmoz(:, z), then
         lstep(2, z) + grdmss(:, z);
  then   the particle array in blocks of 5000 particles of nwpp words:
         x, y, vx, vy (grid units, per-step velocities, time-centred and
         scaled by the zone step factors), the population flag and the
         word offset of the next particle of the same list (-1 ends it).

read_dmp() decodes that into arrays in slot order, with each particle's
list (zone 1..nzones, nzones + 1 off the grid).  write_dmp() writes a dump
in the same layout from a decoded one, with new particles, step and masses
if given: the lists link their particles in slot order, which is how the
reference's own set-up dumps are laid out (a dump read and written back
unchanged is byte-identical).  restart() loads a decoded dump into a
P2DGrid: particles, step counter and zone time stamps, then masset (dumps
are written at full steps, when every zone's mass is current).
"""
from __future__ import annotations

import struct

import numpy as np

DMP_BLOCK = 5000


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


def read_dmp(path) -> dict:
    with open(path, "rb") as f:
        recs = _records(f.read())
    # hedrec's header: two records and one per mass component (hedrec.f),
    # then the step record: irun, istep, noff, ncen(mcmp), nshort, nrsup,
    # suppl, n, numprocs, angoff(3, mcmp), amoff(3, mcmp), rguard: 9 + 7 mcmp
    # words (a component record is 38 words, which that form cannot take)
    h = 3
    while h < len(recs) and not (len(recs[h]) >= 64 and len(recs[h]) % 4 == 0
                                 and (len(recs[h]) // 4 - 9) % 7 == 0):
        h += 1
    if h >= len(recs) - 3:
        raise ValueError("not a dump: no step record")
    r4 = np.frombuffer(recs[h], "<i4")
    mcmp = (r4.size - 9) // 7
    irun, istep, noff = int(r4[0]), int(r4[1]), int(r4[2])
    if r4[mcmp + 5] != 0:
        raise NotImplementedError("dumps with supplementary-force tables")
    lst = np.frombuffer(recs[h + 1], "<i4")
    numprocs, myid, nbod, nwpp = (int(v) for v in lst[:4])
    heads = lst[4:].copy()
    if numprocs != 1 or nwpp < 6:
        raise NotImplementedError("single-process 2-D dumps only")
    nlists = len(heads)
    nz = nlists - 1
    grd1 = np.frombuffer(recs[h + 3], "<f4").copy()
    mesh = grd1.size
    lstep = np.zeros((2, nz), np.int32)
    moz = np.zeros((nz, mesh), np.float32)
    mss = np.zeros((nz, mesh), np.float32)
    mss[0] = grd1
    k = h + 4
    for z in range(1, nz):
        a, b = recs[k], recs[k + 1]
        lstep[0, z] = np.frombuffer(a[:4], "<i4")[0]
        moz[z] = np.frombuffer(a[4:], "<f4")
        lstep[1, z] = np.frombuffer(b[:4], "<i4")[0]
        mss[z] = np.frombuffer(b[4:], "<f4")
        k += 2
    words = b"".join(recs[k:])
    if len(words) != 4 * nbod * nwpp:
        raise ValueError("particle records hold %d bytes, %d expected"
                         % (len(words), 4 * nbod * nwpp))
    pf = np.frombuffer(words, "<f4").reshape(nbod, nwpp).copy()
    pi = np.frombuffer(words, "<i4").reshape(nbod, nwpp).copy()
    iz = np.zeros(nbod, np.int32)
    for li, head in enumerate(heads, start=1):
        loc, seen = int(head), 0
        while loc >= 0:
            if loc % nwpp or seen > nbod:
                raise ValueError("broken particle list %d" % li)
            iz[loc // nwpp] = li
            loc = int(pi[loc // nwpp, nwpp - 1])
            seen += 1
    if np.any(iz == 0):
        raise ValueError("particles on no list")
    return dict(header=recs[:h], steprec=recs[h], rng=recs[h + 2], irun=irun,
                istep=istep, noff=noff, nbod=nbod, nwpp=nwpp, nz=nz,
                lstep=lstep, grdmss=mss, grdmoz=moz,
                x=pf[:, 0].copy(), y=pf[:, 1].copy(), vx=pf[:, 2].copy(),
                vy=pf[:, 3].copy(), iz=iz, flag=pi[:, nwpp - 2].copy(),
                extra=pf[:, 4:nwpp - 2].copy())


def write_dmp(path, d, x=None, y=None, vx=None, vy=None, iz=None, istep=None,
              noff=None, lstep=None, grdmss=None, grdmoz=None):
    """Writes dump d (as read_dmp returns it), with any of its particles,
    step, off-grid count, zone time stamps and mass arrays replaced."""
    x = d["x"] if x is None else x
    y = d["y"] if y is None else y
    vx = d["vx"] if vx is None else vx
    vy = d["vy"] if vy is None else vy
    iz = d["iz"] if iz is None else np.asarray(iz, np.int32)
    n, nwpp, nz = len(x), d["nwpp"], d["nz"]
    flag = d["flag"] if len(d["flag"]) == n else np.ones(n, np.int32)
    extra = d["extra"] if len(d["extra"]) == n else np.zeros((n, nwpp - 6), np.float32)
    istep = d["istep"] if istep is None else int(istep)
    noff = int(np.sum(iz == nz + 1)) if noff is None else int(noff)
    lstep = d["lstep"] if lstep is None else np.asarray(lstep, np.int32)
    mss = d["grdmss"] if grdmss is None else np.asarray(grdmss, np.float32)
    moz = d["grdmoz"] if grdmoz is None else np.asarray(grdmoz, np.float32)
    # lists linked in slot order
    pf = np.zeros((n, nwpp), "<f4")
    pf[:, 0], pf[:, 1], pf[:, 2], pf[:, 3] = x, y, vx, vy
    if nwpp > 6:
        pf[:, 4:nwpp - 2] = extra
    pi = pf.view("<i4")
    pi[:, nwpp - 2] = flag
    heads = np.full(nz + 1, -1, np.int32)
    for li in range(1, nz + 2):
        idx = np.nonzero(iz == li)[0]
        if idx.size:
            heads[li - 1] = idx[0] * nwpp
            nxt = np.append(idx[1:] * nwpp, -1)
            pi[idx, nwpp - 1] = nxt
    step = np.frombuffer(d["steprec"], "<i4").copy()
    step[1], step[2] = istep, noff
    out = [*d["header"], step.tobytes(),
           np.array([1, 0, n, nwpp], "<i4").tobytes() + heads.astype("<i4").tobytes(),
           d["rng"], mss[0].astype("<f4").tobytes()]
    for z in range(1, nz):
        out.append(np.array([lstep[0, z]], "<i4").tobytes() + moz[z].astype("<f4").tobytes())
        out.append(np.array([lstep[1, z]], "<i4").tobytes() + mss[z].astype("<f4").tobytes())
    words = pf.tobytes()
    blk = 4 * nwpp * DMP_BLOCK
    with open(path, "wb") as f:
        for r in out:
            f.write(_record(r))
        for s in range(0, len(words), blk):
            f.write(_record(words[s:s + blk]))


def restart(grid, d):
    """Loads dump d into a P2DGrid: particles and lists, the step counter and
    zone time stamps, then masset.  The velocities in a dump are already
    time-centred: no tmcent follows."""
    grid.restrt(d["x"], d["y"], d["vx"], d["vy"], d["iz"])
    lst = np.zeros((2, d["nz"]), np.int32)
    lst[:, 1:] = d["lstep"][:, 1:]
    grid.set_step(d["istep"], lst)
    grid.masset()
