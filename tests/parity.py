"""Helpers shared by the parity tests: per-bin means, standard errors and the
3-SE / chi^2 comparison of two spectra."""
import math

from scipy import stats


def bin_stats(s, k=0):
    """[(n, mean, se)] per bin for weight slot k."""
    out = []
    for b in range(s.n_bins):
        n = s.count[b]
        if n == 0:
            out.append((0, 0.0, 0.0))
            continue
        m = s.sum_w[k][b] / n
        var = max(s.sum_w2[k][b] / n - m * m, 0.0)
        out.append((n, m, math.sqrt(var / n)))
    return out


def compare_spectra(a, b, k=0, min_count=20):
    """Per-bin z = (m_a - m_b) / sqrt(se_a^2 + se_b^2) over bins where both
    sides have min_count events; returns (zs, chi2, dof, p_value)."""
    zs = []
    for (na, ma, sa), (nb, mb, sb) in zip(bin_stats(a, k), bin_stats(b, k)):
        if na < min_count or nb < min_count:
            continue
        se = math.hypot(sa, sb)
        if se == 0:
            continue
        zs.append((ma - mb) / se)
    chi2 = sum(z * z for z in zs)
    dof = len(zs)
    p = stats.chi2.sf(chi2, dof) if dof else 1.0
    return zs, chi2, dof, p
