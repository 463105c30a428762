"""Ensembles of p2d discs (galhip_ensset) against the same discs run alone.

The reference has no ensemble form (it runs one model per process), so the
parity bar is per-member equivalence: every member of a batched run must be
bit-identical to the same disc in its own galhip_p2dset context -- particles,
zones, fixed-point zone masses, ring ranges, field and potential, off-grid
count, icheck integrals -- in FP32 and FP64, with three time-step zones,
uneven members (one empty, one smaller than a push block) and re-sorts.
The lone-disc path is itself held to the oracle by the p2d parity tests.

HIP graph replay (galhip_set_graph) must give the direct launches' results
bit for bit after every replayed run, with direct work and synchronisation
between the runs.
"""
import numpy as np
import pytest

from galaxy_amd import configs
from galaxy_amd.lib15 import P2DEnsemble, P2DGrid
from tests._model import SMALL

pytestmark = pytest.mark.gpu

# members of a small three-zone ensemble on SMALL's 64 x 128 grid
COUNTS = [3000, 40000, 0, 7, 25000, 1500]
QS = [1.0, 1.5, 1.2, 1.8, 2.0, 1.3]


def _say(*a):
    print("[case]", *a, flush=True)


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    eq = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
    if not np.all(eq):
        bad = ~eq
        d = np.nanmax(np.abs(a.astype(np.float64) - b.astype(np.float64)))
        raise AssertionError("%s: %d of %d differ (max |diff| %.3g), first %s vs %s"
                             % (what, bad.sum(), bad.size, d, a[bad][:4], b[bad][:4]))


def _member(i):
    return configs.Workload("ens-test-m%d" % i, nr=SMALL.nr, na=SMALL.na, hash=SMALL.hash,
                            n=max(COUNTS[i], 1), steps=1, dt=SMALL.dt, nstep=SMALL.nstep,
                            zbr_model=SMALL.zbr_model, Q=QS[i], seed=100 + i)


def _parts(i, rgrid, dt):
    if COUNTS[i] == 0:
        e = np.zeros(0)
        return (e, e, e, e, np.zeros(0, np.int32))
    x, y, vx, vy, iz = configs.particles(_member(i), COUNTS[i], None, rgrid)
    return (x.astype(dt), y.astype(dt), vx.astype(dt), vy.astype(dt), iz)


def _kw():
    kw = configs.grid_kwargs(SMALL, 1)
    kw.pop("pmass")
    return kw


def _pmass():
    return [configs.grid_kwargs(_member(i), max(COUNTS[i], 1))["pmass"]
            for i in range(len(COUNTS))]


def _ensemble(prec, sort_interval=3):
    kw, pm = _kw(), _pmass()
    E = P2DEnsemble(len(COUNTS), pmass=pm, precision=prec, **kw)
    E.greenm()
    E.set_fixforce(*configs.fixforce(SMALL, E.rgrid))
    E.set_fixpot(*configs.fixpot(SMALL, E.rgrid))
    parts = [_parts(i, E.rgrid, E.dtype) for i in range(len(COUNTS))]
    E.restrt(parts)
    E.set_sort_interval(sort_interval)
    return E, parts


def _singles(E, parts, prec, sort_interval=3):
    kw, pm = _kw(), _pmass()
    G = E.get_green()
    out = []
    for i in range(len(COUNTS)):
        S = P2DGrid(precision=prec, pmass=pm[i], **kw)
        S.set_green(G)
        S.set_fixforce(*configs.fixforce(SMALL, S.rgrid))
        S.set_fixpot(*configs.fixpot(SMALL, S.rgrid))
        S.restrt(*parts[i])
        S.set_sort_interval(sort_interval)
        out.append(S)
    return out


def _compare(E, singles, what):
    members = E.unload()
    noff = E.noff
    ints = E.integrals(centred=True)
    for i, S in enumerate(singles):
        tag = "%s member %d" % (what, i)
        o = S.unload()
        for k in ("x", "y", "vx", "vy", "iz"):
            _same(members[i][k], o[k], tag + " " + k)
        for name, a, b in zip(("mfx", "moz", "jrad", "krad", "qbits"), E.get_mass(i),
                              S.get_mass()):
            _same(a, b, tag + " " + name)
        _same(E.get_field(i), S.get_field(), tag + " field")
        assert E.ranges(i) == S.ranges(), tag
        assert noff[i] == S.noff, tag
        if COUNTS[i]:
            si = S.integrals(centred=True)
            for k in ("mass", "cm", "momentum", "ke", "lz", "w_grid", "w_fixed"):
                _same(ints[i][k], si[k], tag + " integral " + k)


def _advance(grids, fn):
    for g in grids:
        fn(g)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_members_equal_lone_discs(prec):
    """masset, findf with the potential, tmcent, 11 steps of run (zones 1,
    2 and 4 due in turn, re-sorts every 3 steps), then findf again: every
    member bit-identical to its lone disc."""
    _say("ensemble vs lone discs", prec, COUNTS)
    E, parts = _ensemble(prec)
    S = _singles(E, parts, prec)
    grids = [E] + S
    _advance(grids, lambda g: g.masset())
    _advance(grids, lambda g: g.findf(True))
    _compare(E, S, "after masset+findf")
    _advance(grids, lambda g: g.tmcent(True))
    _advance(grids, lambda g: g.run(11))
    _advance(grids, lambda g: g.findf(True))
    _compare(E, S, "after 11 steps")
    # single steps through the generic step entry point as well
    _advance(grids, lambda g: (g.findf(False), g.step()))
    _advance(grids, lambda g: g.findf(True))
    _compare(E, S, "after one more step")
    for g in grids:
        g.close()


def test_ensemble_rejects_single_grid_calls():
    E, _ = _ensemble("fp32")
    from galaxy_amd._lib import GalhipError
    with pytest.raises(GalhipError):
        P2DGrid.get_field(E)
    with pytest.raises(GalhipError):
        P2DGrid.integrals(E)
    with pytest.raises(GalhipError):
        P2DGrid.restrt(E, np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3))
    with pytest.raises(GalhipError):
        E.get_field(len(COUNTS))
    E.close()


@pytest.mark.parametrize("kind", ["ensemble", "grid"])
def test_graph_replay_equals_direct_launches(kind):
    """The same run with galhip_set_graph on and off: bit-identical after
    each of three runs (graphs captured in the first, replayed in the
    later ones across zone phases and re-sorts), with direct work between."""
    _say("graph replay", kind)
    if kind == "ensemble":
        A, parts = _ensemble("fp32")
        B, _ = _ensemble("fp32")
    else:
        from tests._model import gpu_model
        A = gpu_model(precision="fp32", n=30000)
        B = gpu_model(precision="fp32", n=30000)
        A.set_sort_interval(3)
        B.set_sort_interval(3)
    A.set_graph(True)
    for g in (A, B):
        g.masset()
        g.findf(True)
        g.tmcent(True)
    for k, nsteps in enumerate((9, 6, 13)):
        A.run(nsteps)
        B.run(nsteps)
        A.sync()
        B.sync()
        if kind == "ensemble":
            ua, ub = A.unload(), B.unload()
            for i in range(len(COUNTS)):
                for key in ("x", "y", "vx", "vy", "iz"):
                    _same(ua[i][key], ub[i][key], "run %d member %d %s" % (k, i, key))
                for name, a, b in zip(("mfx", "moz", "jrad", "krad"), A.get_mass(i)[:4],
                                      B.get_mass(i)[:4]):
                    _same(a, b, "run %d member %d %s" % (k, i, name))
                _same(A.get_field(i), B.get_field(i), "run %d member %d field" % (k, i))
        else:
            ua, ub = A.unload(), B.unload()
            for key in ("x", "y", "vx", "vy", "iz"):
                _same(ua[key], ub[key], "run %d %s" % (k, key))
            for name, a, b in zip(("mfx", "moz", "jrad", "krad"), A.get_mass()[:4],
                                  B.get_mass()[:4]):
                _same(a, b, "run %d %s" % (k, name))
            _same(A.get_field(), B.get_field(), "run %d field" % k)
        # direct work between the replays
        for g in (A, B):
            g.findf(True)
            g.integrals()
    A.close()
    B.close()


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _oracle_fields(orc, parts, lpot=True):
    """The FP64 oracle's field of every non-empty member (its own Green's
    tables, the member's particles in FP64), after masset + findf."""
    kw, pm = _kw(), _pmass()
    out = {}
    for i in range(len(COUNTS)):
        if COUNTS[i] == 0:
            continue
        m = orc.P2D(kw["nr"], kw["na"], kw["hash_"], kw["softl"], tsoft=1, nsect=1,
                    nzones=kw["nzones"], nstep=kw["nstep"], zbr=kw["zbr"], pmass=pm[i],
                    offanal=True, prec="d", use_fx=True)
        x, y, vx, vy, iz = configs.particles(_member(i), COUNTS[i], None, float(m.rgrid))
        m.load(x, y, vx, vy, iz)
        m.masset()
        m.findf(lpot)
        out[i] = m.fld.copy()
    return out


def test_mfma_contraction_vs_fp64_oracle(orc):
    """The split-BF16 MFMA contraction (three bf16 words per operand, six
    cross products, FP32 sums): every member's F_x, F_y and potential within
    1e-5 of the FP64 oracle's field (relative force error, the same bar as
    the FP32 dot products), for the tuned variant and for each variant on
    its own; and within 1e-5 of the FP32 dot products."""
    _say("mfma contraction vs the FP64 oracle")
    E, parts = _ensemble("fp32")
    ref = _oracle_fields(orc, parts)
    E.masset()
    E.findf(True)
    for v in (None, 0, 1, 2):
        V, _ = _ensemble("fp32")
        if v is None:
            V.set_contraction("mfma")
        else:
            V.set_contraction("mfma", variant=v)
        V.masset()
        V.findf(True)
        mode, var, us = V.get_contraction()
        assert mode == 1 and (var == v if v is not None else var in (0, 1, 2))
        if v is None:
            _say("mfma tuned variant %d, variant times (us) %s" % (var, us))
            assert np.all(us > 0)
        worst = worst_dot = 0.0
        for i, f64 in ref.items():
            fv, fe = V.get_field(i), E.get_field(i)
            for t in range(3):
                e = _rel(fv[t], f64[t])
                assert e <= 1e-5, (v, i, t, e)
                worst = max(worst, e)
                worst_dot = max(worst_dot, _rel(fv[t], fe[t]))
        dot = max(_rel(E.get_field(i)[t], f64[t]) for i, f64 in ref.items() for t in range(3))
        _say("variant %s: force error vs FP64 oracle %.2e (FP32 dot products %.2e), "
             "vs the dot products %.2e" % (var, worst, dot, worst_dot))
        assert worst_dot <= 1e-5
        V.close()
    E.close()


def test_mfma_contraction_graph_and_lone_grid():
    """A graph-replayed run with the MFMA contraction equals the direct run
    bit for bit; a lone grid with the MFMA contraction is within 1e-5 of the
    FP32 dot products."""
    _say("mfma contraction: graphs, lone grid")
    M, _ = _ensemble("fp32")
    M.set_contraction("mfma")
    M.masset()
    M.findf(True)
    mode, var, us = M.get_contraction()
    # graph replay with the GEMM in the step (the variant M tuned to)
    A, _ = _ensemble("fp32")
    A.set_contraction("mfma", variant=var)
    A.set_graph(True)
    for g in (A, M):
        g.tmcent(True) if g is M else (g.masset(), g.findf(True), g.tmcent(True))
    for k in range(2):
        A.run(7)
        M.run(7)
        ua, um = A.unload(), M.unload()
        for i in range(len(COUNTS)):
            for key in ("x", "y", "vx", "vy", "iz"):
                _same(ua[i][key], um[i][key], "mfma graph run %d member %d %s" % (k, i, key))
    for g in (M, A):
        g.close()
    # a lone grid
    from tests._model import gpu_model
    P = gpu_model(precision="fp32", n=30000)
    Q = gpu_model(precision="fp32", n=30000)
    Q.set_contraction("mfma")
    for g in (P, Q):
        g.masset()
        g.findf(True)
    fp, fq = P.get_field(), Q.get_field()
    for t in range(3):
        assert _rel(fq[t], fp[t]) < 1e-5, (t, _rel(fq[t], fp[t]))
    P.close()
    Q.close()
