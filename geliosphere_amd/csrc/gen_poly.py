"""Chebyshev fits (near-minimax) of the polynomial kernels in gs_math.hpp,
computed at 60 digits with mpmath.  Prints the coefficients and the worst
relative error of each fit on its interval.

  log:  ln((1+s)/(1-s)) = 2s + 2s w R(w),  w = s^2 in [0, (3-2sqrt2)^2]
  sin:  sin a = a - a w S(w),              w = a^2 in [0, (pi/4)^2]
  cos:  cos a = 1 - w/2 + w^2 C(w)
  atan: atan t = t - t w A(w),            w = t^2 in [0, tan(pi/16)^2]
"""
import mpmath as mp

mp.mp.dps = 60
S_MAX = 3 - 2 * mp.sqrt(2)          # |s| for m in [1/sqrt2, sqrt2]
A_MAX = mp.pi / 4


def fit(f, a, b, deg):
    coeffs, err = mp.chebyfit(f, [a, b], deg + 1, error=True)
    return coeffs[::-1], err        # ascending powers


def R(w):
    if w == 0:
        return mp.mpf(1) / 3
    s = mp.sqrt(w)
    return (mp.log((1 + s) / (1 - s)) - 2 * s) / (2 * s * w)


def S(w):
    if w == 0:
        return mp.mpf(1) / 6
    a = mp.sqrt(w)
    return (a - mp.sin(a)) / (a * w)


def C(w):
    if w == 0:
        return mp.mpf(1) / 24
    a = mp.sqrt(w)
    return (mp.cos(a) - 1 + w / 2) / (w * w)


T_MAX = mp.tan(mp.pi / 16)


def A(w):
    if w == 0:
        return mp.mpf(1) / 3
    t = mp.sqrt(w)
    return (t - mp.atan(t)) / (t * w)


def worst(kind, coeffs, n=4000):
    """Worst relative error of the final function value over the interval."""
    e = mp.mpf(0)
    for i in range(1, n + 1):
        if kind == "log":
            s = S_MAX * i / n
            w = s * s
            p = sum(c * w ** k for k, c in enumerate(coeffs))
            v, ref = 2 * s + 2 * s * w * p, mp.log((1 + s) / (1 - s))
        elif kind == "sin":
            a = A_MAX * i / n
            w = a * a
            p = sum(c * w ** k for k, c in enumerate(coeffs))
            v, ref = a - a * w * p, mp.sin(a)
        elif kind == "atan":
            t = T_MAX * i / n
            w = t * t
            p = sum(c * w ** k for k, c in enumerate(coeffs))
            v, ref = t - t * w * p, mp.atan(t)
        else:
            a = A_MAX * i / n
            w = a * a
            p = sum(c * w ** k for k, c in enumerate(coeffs))
            v, ref = 1 - w / 2 + w * w * p, mp.cos(a)
        e = max(e, abs(v - ref) / abs(ref))
    return e


if __name__ == "__main__":
    for kind, f, hi, deg in (("log", R, S_MAX ** 2, 6), ("sin", S, A_MAX ** 2, 5), ("cos", C, A_MAX ** 2, 5),
                             ("atan", A, T_MAX ** 2, 8)):
        c, _ = fit(f, 0, hi, deg)
        print(kind, "degree", deg, "worst rel err %.3e" % float(worst(kind, c)))
        for k, v in enumerate(c):
            print("   c%d = %s" % (k, mp.nstr(v, 22)))
