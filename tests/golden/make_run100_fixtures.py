"""Writes tests/golden/run100_dmpNN.npz from the reference's shipped restart
dumps EXAMPLES/p2d/run100.dmp{0,10,80} (the cold Kuzmin-Toomre disc standard
test, EXAMPLES/p2d/run100.dat).  Run in a container where the reference tree
exists; the fixtures it writes are data only (inputs and expected outputs).

Dump layout (SRC/lib15/dumpp.f:57-198): Fortran unformatted sequential
records with 4-byte markers; record 3 holds irun, istep; record 4 holds
numprocs, myid, nbod, nwpp and the list heads islist(1, 1:nlists); records
6-8 hold grdmss(:,1) (scaled), (lstep(1,2), grdmoz(:,2)) and
(lstep(2,2), grdmss(:,2)); particle blocks follow, nwpp = 6 words per
particle (x, y, vx, vy, population flag, next-pointer as integers).
"""
import os
import struct
import sys

import numpy as np

SRC = "/root/reference/EXAMPLES/p2d"
OUT = os.path.dirname(os.path.abspath(__file__))


def records(path):
    b = open(path, "rb").read()
    o, out = 0, []
    while o < len(b):
        n = struct.unpack("<i", b[o:o + 4])[0]
        out.append(b[o + 4:o + 4 + n])
        o += 8 + n
    return out


def convert(tag):
    r = records(os.path.join(SRC, "run100." + tag))
    g1 = np.frombuffer(r[1], "<i4")
    nr, na, ng = int(g1[0]), int(g1[1]), int(g1[2])
    rgrid = float(np.frombuffer(r[1], "<f4")[3])
    h0f = np.frombuffer(r[0], "<f4")
    lscale, ts, pmass, softl = (float(h0f[5]), float(h0f[6]),
                                float(h0f[11]), float(h0f[12]))
    irun, istep = (int(v) for v in np.frombuffer(r[3], "<i4")[:2])
    rec4 = np.frombuffer(r[4], "<i4")
    nbod, nwpp = int(rec4[2]), int(rec4[3])
    heads = rec4[4:]
    pf = np.concatenate([np.frombuffer(x, "<f4") for x in r[9:]]).reshape(-1, nwpp)
    pi = np.concatenate([np.frombuffer(x, "<i4") for x in r[9:]]).reshape(-1, nwpp)
    assert pf.shape[0] == nbod
    iz = np.zeros(nbod, np.int8)
    for lst, head in enumerate(heads, start=1):
        loc = int(head)
        while loc >= 0:
            iz[loc // nwpp] = lst
            loc = int(pi[loc // nwpp, nwpp - 1])
    assert np.all(iz > 0)
    grd1 = np.frombuffer(r[6], "<f4").copy()
    lstep1 = int(np.frombuffer(r[7][:4], "<i4")[0])
    moz2 = np.frombuffer(r[7][4:], "<f4").copy()
    lstep2 = int(np.frombuffer(r[8][:4], "<i4")[0])
    grd2 = np.frombuffer(r[8][4:], "<f4").copy()
    np.savez_compressed(
        os.path.join(OUT, "run100_%s.npz" % tag),
        x=pf[:, 0].copy(), y=pf[:, 1].copy(), vx=pf[:, 2].copy(),
        vy=pf[:, 3].copy(), iz=iz, grdmss1=grd1, grdmoz2=moz2, grdmss2=grd2,
        lstep=np.array([lstep1, lstep2]),
        scalars=np.array([nr, na, ng, istep, nbod]),
        reals=np.array([rgrid, lscale, ts, pmass, softl]))
    print(tag, "istep", istep, "zones", np.bincount(iz)[1:], "ng", ng)


if __name__ == "__main__":
    for t in (sys.argv[1:] or ["dmp0", "dmp10", "dmp80"]):
        convert(t)
