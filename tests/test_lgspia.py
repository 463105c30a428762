"""Log-spiral coefficients (lgspia, SRC/lib15/lgspia.f:44-110, normalised
as phyprp.f:182-197).

The oracle's restatement (oracle.lgspia: the reference's complex
recurrences) is pinned by what the definition fixes:
  - the closed form A[j, im] = mean exp(i im (nsect theta + k_j a)),
    a = nsect ln(R / lscale) / 4, k_j = j - (np // 2 + 1) from the centre
    up and, below it, the pitch the reference's mirror index holds
    (k_j = j + np // 2 + 1 - np - 1: the same for odd np; for even np
    index np / 2 stays empty), evaluated directly with exp and arctan2 (no
    recurrence), for odd and even np, one to three sectors, off-grid
    particles with and without offanal;
  - particles on one two-armed logarithmic spiral give modulus 1 at the
    m = 2 coefficient of its pitch index, the largest of all pitches, a
    small value at the opposite pitch and 0 for m = 1.
The reference's own values are checked in tests/test_reference_binary.py;
the GPU's against the oracle in tests/test_gpu_lgspia.py.
"""
import numpy as np
import pytest


def _closed(x, y, iz, nz, offanal, nsect, np_, nm, lscale):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    use = (iz >= 1) & ((iz <= nz) | (offanal & (iz == nz + 1)))
    R = np.hypot(x, y)
    use &= R > 0
    th = np.arctan2(y[use], x[use])
    a = 0.25 * nsect * np.log(R[use] / lscale)
    jc = np_ // 2 + 1
    A = np.zeros((np_, nm), np.complex128)
    for j in range(1, np_ + 1):
        # pitch of index j: j - jc from the centre up; below it the mirror
        # np + 1 - j of index jc + k holds -k, which for even np puts pitch
        # j - np/2 at j and leaves j = np/2 empty
        if j >= jc:
            k = j - jc
        elif np_ % 2 == 0 and j == np_ // 2:
            continue
        else:
            k = j + jc - np_ - 1
        for im in range(1, nm + 1):
            A[j - 1, im - 1] = np.exp(1j * im * (nsect * th + k * a)).sum()
    A /= max(use.sum(), 1)
    return A.real.reshape(-1), A.imag.reshape(-1)


@pytest.mark.parametrize("nsect,np_,nm,offanal", [(1, 61, 4, True), (2, 60, 3, False),
                                                   (2, 7, 1, True), (3, 1, 2, True)])
def test_oracle_matches_closed_form(orc, nsect, np_, nm, offanal):
    rng = np.random.default_rng(nsect * 100 + np_)
    n = 3000
    R = np.exp(rng.uniform(-2, 2.5, n))
    ph = rng.uniform(-np.pi, np.pi, n)
    ls = 7.5
    x, y = R * ls * np.cos(ph), R * ls * np.sin(ph)
    iz = rng.integers(1, 4, n).astype(np.int32)          # two zones + off-grid list 3
    x[:3] = 0.0
    y[:3] = 0.0                                         # R = 0: left out
    c, s = orc.lgspia(x, y, iz, 2, offanal, nsect, np_, nm, ls)
    cc, ss = _closed(x, y, iz, 2, offanal, nsect, np_, nm, ls)
    assert np.abs(c - cc).max() < 1e-13 and np.abs(s - ss).max() < 1e-13


def test_oracle_logarithmic_spiral(orc):
    np_, nm, ls = 21, 2, 1.0
    jc = np_ // 2 + 1
    ksel = 4
    R = np.repeat(np.exp(np.linspace(-3.0, 4.0, 2000)), 2)
    a = 0.25 * np.log(R)
    # 2 (theta + ksel a) is the same for every particle on either arm (the
    # arms pi apart, one particle of each at every radius)
    th = -ksel * a + np.pi * (np.arange(R.size) % 2)
    x, y = R * np.cos(th), R * np.sin(th)
    c, s = orc.lgspia(x, y, np.ones(R.size, np.int32), 1, True, 1, np_, nm, ls)
    A = np.abs((c + 1j * s).reshape(np_, nm))
    j = jc + ksel
    assert abs(A[j - 1, 1] - 1.0) < 1e-12
    assert int(np.argmax(A[:, 1])) == j - 1
    assert A[jc - ksel - 1, 1] < 0.1          # the opposite pitch
    assert A[:, 0].max() < 1e-12               # m = 1 cancels between the arms
