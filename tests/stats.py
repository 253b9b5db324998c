"""Two-sample statistics for comparing Monte Carlo histograms."""
import numpy as np


def two_sample_z(c1, n1, c2, n2, min_count=10, min_small=0.0):
    """Per-bin z of the difference of two binomial proportions (pooled variance),
    for bins with at least `min_count` counts in total and at least `min_small`
    expected counts in the smaller sample (with very unequal samples a bin the
    small one expects ~1 count of has Poisson, not Gaussian, tails)."""
    c1 = np.asarray(c1, dtype=np.float64).ravel()
    c2 = np.asarray(c2, dtype=np.float64).ravel()
    tot = c1 + c2
    sel = (tot >= min_count) & (tot * (min(n1, n2) / (n1 + n2)) >= min_small)
    p = tot[sel] / (n1 + n2)
    var = p * (1 - p) * (1.0 / n1 + 1.0 / n2)
    z = (c1[sel] / n1 - c2[sel] / n2) / np.sqrt(var)
    return z


def assert_same_distribution(c1, n1, c2, n2, what="", min_bins=20, min_small=0.0):
    """Every bin within 3 standard errors up to the expected 0.27 % (allowing 1 %),
    no bin beyond 6, and mean z^2 near 1 (a biased kernel shifts it)."""
    z = two_sample_z(c1, n1, c2, n2, min_small=min_small)
    assert z.size >= min_bins, f"{what}: only {z.size} populated bins"
    frac3 = float(np.mean(np.abs(z) > 3))
    zmax = float(np.max(np.abs(z)))
    chi = float(np.mean(z ** 2))
    assert frac3 <= 0.01 + 3.0 / z.size, f"{what}: {frac3:.4f} of {z.size} bins beyond 3 sigma"
    if zmax >= 6.0:
        a1 = np.asarray(c1, dtype=np.float64).ravel()
        a2 = np.asarray(c2, dtype=np.float64).ravel()
        tot = a1 + a2
        idx = np.flatnonzero((tot >= 10) & (tot * (min(n1, n2) / (n1 + n2)) >= min_small))[int(np.argmax(np.abs(z)))]
        raise AssertionError(f"{what}: max |z| = {zmax:.2f} at bin {idx}: counts {a1[idx]:.0f} of {n1} vs "
                             f"{a2[idx]:.0f} of {n2} ({z.size} bins, frac3 {frac3:.4f}, mean z^2 {chi:.3f})")
    assert 0.7 < chi < 1.35, f"{what}: mean z^2 = {chi:.3f} over {z.size} bins"
    return dict(bins=int(z.size), frac3=frac3, zmax=zmax, chi2=chi)
