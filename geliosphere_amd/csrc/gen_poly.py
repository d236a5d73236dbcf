"""Chebyshev fits (near-minimax) of the polynomial kernels in gs_math.hpp,
computed at 60 digits with mpmath.  Prints the coefficients and the worst
relative error of each fit on its interval.

  log:  ln((1+s)/(1-s)) = 2s + 2s w R(w),  w = s^2 in [0, (3-2sqrt2)^2]
  atan: atan t = t - t w A(w),            w = t^2 in [0, tan(pi/16)^2]
  log1p: ln(1+r) = r + r^2 L(r),           r in [-R_LOG, R_LOG]

With --tables, also writes gs_bm_tables.h: the Box-Muller tables of the
table-driven ln and sin/cos (Tang's method: a table entry near the
argument plus a short polynomial in the remainder).
  trig:  (sin, cos)(2 pi j / 256), j = 0..255
  log:   per interval i of the log reduction (gs_math.hpp, m2log_u32):
         invc ~ 1/centre (1 exactly for the interval around 1) and
         -ln(invc) as a double-double (hi, lo), all three times -2 (exact),
         padded to 4 doubles
"""
import sys
import mpmath as mp

mp.mp.dps = 60
S_MAX = 3 - 2 * mp.sqrt(2)          # |s| for m in [1/sqrt2, sqrt2]


def fit(f, a, b, deg):
    coeffs, err = mp.chebyfit(f, [a, b], deg + 1, error=True)
    return coeffs[::-1], err        # ascending powers


def R(w):
    if w == 0:
        return mp.mpf(1) / 3
    s = mp.sqrt(w)
    return (mp.log((1 + s) / (1 - s)) - 2 * s) / (2 * s * w)


T_MAX = mp.tan(mp.pi / 16)


def A(w):
    if w == 0:
        return mp.mpf(1) / 3
    t = mp.sqrt(w)
    return (t - mp.atan(t)) / (t * w)


R_LOG = mp.mpf("0.0040")            # > 1/256, the largest |r| of the log reduction
LOG_OFF = 0x3fe5f00000000000        # bits of 0.685546875: 1.0 is the centre of interval 80
LOG_BITS = 7


def L(r):
    if r == 0:
        return mp.mpf(-1) / 2
    return (mp.log1p(r) - r) / (r * r)


def dbl_hex(v):
    return float(v).hex()


def log_intervals():
    """(lo, hi) of the z range of each interval: z = ix - (e << 52) with
    tmp = ix - LOG_OFF, e = tmp >> 52, i = (tmp >> 45) & 127."""
    import struct
    out = []
    for i in range(1 << LOG_BITS):
        lo = LOG_OFF + (i << (52 - LOG_BITS))
        hi = lo + (1 << (52 - LOG_BITS))
        d = lambda u: mp.mpf(struct.unpack("<d", struct.pack("<Q", u))[0])
        out.append((d(lo), d(hi)))
    return out


def write_tables(path):
    trig = []
    for j in range(256):
        a = 2 * mp.pi * j / 256
        trig += [mp.sin(a), mp.cos(a)]
    logt = []
    worst_r = mp.mpf(0)
    for lo, hi in log_intervals():
        if lo <= 1 < hi:
            invc = mp.mpf(1)
        else:
            invc = mp.mpf(float(2 / (lo + hi)))
        worst_r = max(worst_r, abs(lo * invc - 1), abs(hi * invc - 1))
        lc = -mp.log(invc)
        h = mp.mpf(float(lc))
        # stored times -2 (exact): the kernels form -2 ln u directly
        logt += [-2 * invc, -2 * h, -2 * mp.mpf(float(lc - h)), mp.mpf(0)]
    assert worst_r < R_LOG, worst_r
    with open(path, "w") as f:
        f.write("// gs_bm_tables.h - written by gen_poly.py --tables; do not edit.\n")
        f.write("// Box-Muller tables of gs_math.hpp (sin/cos of 2 pi j/256; ln reduction\n")
        f.write("// entries -2 invc, 2 ln(invc) hi, lo, 0).  Largest |r| of the log reduction %s.\n"
                % mp.nstr(worst_r, 6))
        f.write("#pragma once\n\n")
        for name, vals in (("GS_BM_TRIG_TABLE", trig), ("GS_BM_LOG_TABLE", logt)):
            f.write("#define %s \\\n    {" % name)
            for k, v in enumerate(vals):
                if k and k % 4 == 0:
                    f.write(" \\\n     ")
                f.write(" %s," % dbl_hex(v))
            f.write(" }\n\n")


def worst(kind, coeffs, n=4000):
    """Worst relative error of the final function value over the interval."""
    e = mp.mpf(0)
    for i in range(1, n + 1):
        if kind == "log":
            s = S_MAX * i / n
            w = s * s
            p = sum(c * w ** k for k, c in enumerate(coeffs))
            v, ref = 2 * s + 2 * s * w * p, mp.log((1 + s) / (1 - s))
        elif kind == "log1p":
            r = R_LOG * (2 * i - n) / n
            if r == 0:
                continue
            v, ref = r + r * r * sum(c * r ** k for k, c in enumerate(coeffs)), mp.log1p(r)
        elif kind == "atan":
            t = T_MAX * i / n
            w = t * t
            p = sum(c * w ** k for k, c in enumerate(coeffs))
            v, ref = t - t * w * p, mp.atan(t)
        e = max(e, abs(v - ref) / abs(ref))
    return e


if __name__ == "__main__":
    if "--tables" in sys.argv:
        import os
        write_tables(os.path.join(os.path.dirname(os.path.abspath(__file__)), "gs_bm_tables.h"))
    for kind, f, lo, hi, deg in (("log", R, 0, S_MAX ** 2, 6), ("atan", A, 0, T_MAX ** 2, 7),
                                 ("log1p", L, -R_LOG, R_LOG, 4)):
        c, _ = fit(f, lo, hi, deg)
        print(kind, "degree", deg, "worst rel err %.3e" % float(worst(kind, c)))
        for k, v in enumerate(c):
            print("   c%d = %s" % (k, mp.nstr(v, 22)))
