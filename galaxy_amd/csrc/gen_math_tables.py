"""Generates the constant tables of galhip_math.h with 60-digit decimal
arithmetic (python galaxy_amd/csrc/gen_math_tables.py > tables.txt).

  LN2_HI + LN2_LO            ln 2, hi with 21 trailing zero bits so that
                             e * LN2_HI is exact for |e| < 2^20
  LOGC_HI[j] + LOGC_LO[j]    ln(1 + j/32), j = 0..31, hi likewise truncated
                             to 32 significant bits (hi sums stay exact)
  ATANC_HI[j] + ATANC_LO[j]  atan(j/16), j = 0..16
  PIO2_HI/LO, PI_HI/LO       pi/2, pi
"""
from decimal import Decimal, getcontext

getcontext().prec = 60


def ln(x):
    return Decimal(x).ln()


def atan(x):
    # atan by the argument-halving identity and the Taylor series
    x = Decimal(x)
    k = 0
    while abs(x) > Decimal("0.05"):
        x = x / (1 + (1 + x * x).sqrt())
        k += 1
    s, term, n = Decimal(0), x, 1
    x2 = x * x
    while True:
        t = term / n
        if abs(t) < Decimal(10) ** -55:
            break
        s += t
        term *= -x2
        n += 2
    return s * (2 ** k)


def pi():
    return 4 * atan(1)


def split(v, bits):
    """hi = v rounded to `bits` significant bits (as a double), lo = v - hi."""
    if v == 0:
        return 0.0, 0.0
    import math
    f = float(v)
    e = math.frexp(f)[1]
    q = Decimal(2) ** (bits - e)
    hi = float((v * q).to_integral_value() / q)
    lo = float(v - Decimal(hi))
    return hi, lo


def fmt(x):
    return float.hex(float(x))


def main():
    hi, lo = split(ln(2), 32)
    print("LN2_HI %s LN2_LO %s" % (fmt(hi), fmt(lo)))
    for j in range(32):
        h, l = split(ln(1 + Decimal(j) / 32), 32)
        print("LOGC %d %s %s" % (j, fmt(h), fmt(l)))
    for j in range(17):
        h, l = split(atan(Decimal(j) / 16), 53)
        print("ATANC %d %s %s" % (j, fmt(h), fmt(l)))
    for name, v in (("PIO2", pi() / 2), ("PI", pi())):
        h, l = split(v, 53)
        print("%s %s %s" % (name, fmt(h), fmt(l)))


if __name__ == "__main__":
    main()
