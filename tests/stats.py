"""Two-sample statistics for comparing Monte Carlo histograms."""
import numpy as np


def two_sample_z(c1, n1, c2, n2, min_count=10):
    """Per-bin z of the difference of two binomial proportions (pooled variance),
    for bins with at least `min_count` counts in total."""
    c1 = np.asarray(c1, dtype=np.float64).ravel()
    c2 = np.asarray(c2, dtype=np.float64).ravel()
    tot = c1 + c2
    sel = tot >= min_count
    p = tot[sel] / (n1 + n2)
    var = p * (1 - p) * (1.0 / n1 + 1.0 / n2)
    z = (c1[sel] / n1 - c2[sel] / n2) / np.sqrt(var)
    return z


def assert_same_distribution(c1, n1, c2, n2, what="", min_bins=20):
    """Every bin within 3 standard errors up to the expected 0.27 % (allowing 1 %),
    no bin beyond 6, and mean z^2 near 1 (a biased kernel shifts it)."""
    z = two_sample_z(c1, n1, c2, n2)
    assert z.size >= min_bins, f"{what}: only {z.size} populated bins"
    frac3 = float(np.mean(np.abs(z) > 3))
    zmax = float(np.max(np.abs(z)))
    chi = float(np.mean(z ** 2))
    assert frac3 <= 0.01 + 3.0 / z.size, f"{what}: {frac3:.4f} of {z.size} bins beyond 3 sigma"
    assert zmax < 6.0, f"{what}: max |z| = {zmax:.2f}"
    assert 0.7 < chi < 1.35, f"{what}: mean z^2 = {chi:.3f} over {z.size} bins"
    return dict(bins=int(z.size), frac3=frac3, zmax=zmax, chi2=chi)
