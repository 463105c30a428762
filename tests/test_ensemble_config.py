"""The ensemble workload (64 discs, log-uniform counts in [1e5, 4e6], Q in
[1, 2], one seed per member) and its particle generator, on the CPU."""
import math

import numpy as np

from galaxy_amd import configs


def test_member_draws():
    w = configs.C4
    counts, Q = w.members()
    assert len(counts) == w.nmem == 64
    assert counts.min() >= 1e5 and counts.max() <= 4e6
    assert Q.min() >= 1.0 and Q.max() <= 2.0
    assert w.n == counts.sum()
    # log-uniform: the mean of log10 counts near the middle of [5, log10 4e6]
    mid = 0.5 * (5.0 + math.log10(4e6))
    assert abs(np.log10(counts).mean() - mid) < 0.25
    c2, Q2 = w.members()
    assert np.array_equal(counts, c2) and np.array_equal(Q, Q2)
    # a scaled ensemble keeps the shape
    cs, _ = w.members(0.01)
    assert np.all(np.abs(cs - np.rint(0.01 * 10 ** np.log10(counts))) <= 1)


def test_members_are_distinct_discs():
    w = configs.C4
    counts, Q = w.members()
    m3, m7 = w.member(3, 1000), w.member(7, 1000)
    assert m3.seed != m7.seed and m3.Q == Q[3] and m7.Q == Q[7]
    assert (m3.nr, m3.na, m3.hash, m3.dt) == (w.nr, w.na, w.hash, w.dt)
    rg = w.rgrid_d()
    p3 = configs.particles(m3, 1000, None, rg)
    p7 = configs.particles(m7, 1000, None, rg)
    assert not np.array_equal(p3[0], p7[0])


def test_parallel_generation_matches_serial():
    w = configs.C4
    counts = np.array([3000, 0, 50, 2000])
    rg = w.rgrid_d()
    serial = [configs.particles(w.member(i, n), int(n), None, rg) if n else None
              for i, n in enumerate(counts)]
    par = configs.ensemble_particles(w, counts, rg, workers=2, min_parallel=0)
    for s, p, n in zip(serial, par, counts):
        assert len(p[0]) == n
        if s is None:
            continue
        for a, b in zip(s, p):
            assert np.array_equal(a, b)
