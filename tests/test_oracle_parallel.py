"""The oracle's multi-core particle loop (orc_push_par_*, OpenMP), which
the CPU baseline times on every host core, against its serial loop: eight
steps through the three-zone cycle, both precisions, positions,
velocities, zones, fixed-point mass sums, ring ranges and the off-grid
count identical (the integer sums commute, so the thread split cannot
change them)."""
import numpy as np
import pytest

from tests._model import SMALL, oracle_model


@pytest.mark.parametrize("prec", ["f", "d"])
def test_parallel_push_equals_serial(orc, prec):
    a, parts = oracle_model(orc, SMALL, prec=prec)
    b, _ = oracle_model(orc, SMALL, prec=prec, G=a.G)
    a.masset()
    b.masset()
    for k in range(8):
        a.findf()
        b.findf()
        a.step()
        b.step(parallel=True)
        for key in ("x", "y", "vx", "vy", "iz", "mfx", "mozfx", "jrad", "krad", "moved"):
            assert np.array_equal(getattr(a, key), getattr(b, key)), (k, key)
        assert a.noff == b.noff and a.mzone == b.mzone
