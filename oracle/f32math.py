"""FP32 exp / log / log1p / expm1 of the oracle -- TEST INFRASTRUCTURE ONLY.

The same operation sequence as the kernels' pgmax_amd/csrc/pgx_math.h, in
NumPy float32 (every operation rounds once; NumPy does not fuse, and the
kernels' fused multiply-adds are computed here by ``fma``, which rounds the
exact a * b + c once), so the oracle's FP32 restatement and the GPU agree bit
for bit on transcendentals. The constructions are textbook: Cody-Waite
reduction with a degree-6 polynomial 1 + r + r^2 q(r) (Horner form in fused
multiply-adds; q fitted for the least maximum relative error on |r| <= ln2/2
with the first two terms held at 1, coefficients rounded to FP32) for exp,
the atanh series for log, and the exact-argument
corrections log(u) * x / (u - 1) and (u - 1) * x / log(u) for log1p and
expm1. tests/test_oracle.py checks every function against float64 (<= 2.5
ulp) and ``fma`` against exact rational arithmetic.
"""

import numpy as np

F = np.float32
LN2_HI = F(0.693359375)
LN2_LO = F(-2.12194440054690583e-4)
INV_LN2 = F(1.44269504088896341)
EXP_HI = F(88.7228317)
EXP_LO = F(-103.972084)
SQRT_HALF = F(0.707106781186547524)
ONE, TWO, HALF = F(1), F(2), F(0.5)
# e^r = 1 + r + r^2 (c2 + c3 r + c4 r^2 + c5 r^3 + c6 r^4), highest first
EXP_C = tuple(F(float.fromhex(h)) for h in
              ("0x1.6c07b2p-10", "0x1.12609cp-7", "0x1.555748p-5", "0x1.555452p-3",
               "0x1.fffffcp-2"))
D = {n: F(2) / F(n) for n in (3, 5, 7, 9, 11)}


def _a(x):
  return np.asarray(x, dtype=F)


def fma(a, b, c):
  """float32 fused multiply-add: a * b + c rounded once to nearest even.

  The product of two float32 is exact in float64; the float64 sum s and its
  exact error e (Knuth's two-sum) give a + b + c = s + e exactly. Rounding s
  to float32 is the correct result unless s lies exactly halfway between two
  float32 values and e is not zero -- then the exact value lies on the side
  of the halfway point that e points to. (A halfway point is a float64
  number, so s cannot lie strictly on the other side of one from s + e.)
  """
  a64 = _a(a).astype(np.float64)
  b64 = _a(b).astype(np.float64)
  c64 = _a(c).astype(np.float64)
  with np.errstate(all="ignore"):
    p = a64 * b64
    s = p + c64
    bv = s - c64
    e = (p - bv) + (c64 - (s - bv))
    r = s.astype(F)
    d = s - r.astype(np.float64)
    up = np.nextafter(r, np.where(d > 0, F(np.inf), F(-np.inf)).astype(F))
    half = (up.astype(np.float64) - r.astype(np.float64)) * 0.5
    tie = np.isfinite(s) & np.isfinite(up) & (d != 0) & (d == half) & (e * d > 0)
    return np.where(tie, up, r).astype(F)


def exp(x):
  x = _a(x)
  with np.errstate(all="ignore"):
    xc = np.where(np.isfinite(x), np.clip(x, EXP_LO, EXP_HI), F(0))
    k = np.rint(xc * INV_LN2)
    r = fma(-k, LN2_HI, xc)
    r = fma(-k, LN2_LO, r)
    p = EXP_C[0]
    for c in EXP_C[1:] + (ONE, ONE):
      p = fma(p, r, c)
    res = np.ldexp(p, k.astype(np.int32)).astype(F)
    res = np.where(x < EXP_LO, F(0), res)
    res = np.where(x > EXP_HI, F(np.inf), res)
    res = np.where(np.isnan(x), x, res)
  return res.astype(F)


def exp_shift(x):
  """The kernels' select-free exp for arguments <= 0 or -inf (never NaN): the
  core evaluated at max(x, -104). Equal to exp on that domain, which
  tests/test_oracle.py checks; the oracle itself always calls exp."""
  x = _a(x)
  with np.errstate(all="ignore"):
    xc = np.maximum(x, F(-104.0))
    k = np.rint(xc * INV_LN2)
    r = fma(-k, LN2_HI, xc)
    r = fma(-k, LN2_LO, r)
    p = EXP_C[0]
    for c in EXP_C[1:] + (ONE, ONE):
      p = fma(p, r, c)
    return np.ldexp(p, k.astype(np.int32)).astype(F)


def log(x):
  x = _a(x)
  with np.errstate(all="ignore"):
    m, e = np.frexp(x)
    m = m.astype(F)
    small = m < SQRT_HALF
    m = np.where(small, m * TWO, m)
    e = np.where(small, e - 1, e)
    f = m - ONE
    s = f / (TWO + f)
    z = s * s
    q = D[11]
    q = q * z + D[9]
    q = q * z + D[7]
    q = q * z + D[5]
    q = q * z + D[3]
    logm = f - s * (f - q * z)
    ef = e.astype(F)
    res = (ef * LN2_LO + logm) + ef * LN2_HI
    res = np.where(x == F(np.inf), x, res)
    res = np.where(x == F(0), F(-np.inf), res)
    res = np.where(x < F(0), F(np.nan), res)
    res = np.where(np.isnan(x), x, res)
  return res.astype(F)


def log1p(x):
  x = _a(x)
  with np.errstate(all="ignore"):
    u = ONE + x
    res = log(u) * (x / (u - ONE))
    res = np.where(u == ONE, x, res)
    res = np.where(x == F(np.inf), x, res)
    res = np.where(np.isnan(x), x, res)
  return res.astype(F)


def expm1(x):
  x = _a(x)
  with np.errstate(all="ignore"):
    u = exp(x)
    um1 = u - ONE
    res = um1 * (x / log(u))
    res = np.where(u == F(np.inf), u, res)
    res = np.where(um1 == F(-1), F(-1), res)
    res = np.where(u == ONE, x, res)
    res = np.where(np.isnan(x), x, res)
  return res.astype(F)
