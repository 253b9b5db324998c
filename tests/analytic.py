"""Literature solutions used as statistical oracles (the reference's +Comparison
package restated): Paasschens (1997) exact 2-D solution of the radiative
transfer equation for isotropic scattering, the form of
+Comparison/analyticalPaasschens.m:65-68 without the ballistic delta term."""
import numpy as np


def paasschens_2d_diffuse(r, t, Sigma, v):
    """Multiply-scattered energy density per unit area at distance r and time t
    for a unit isotropic point source (2-D, no absorption)."""
    a = Sigma * np.asarray(t, dtype=np.float64)
    b = np.asarray(r, dtype=np.float64) * Sigma / v
    a, b = np.broadcast_arrays(a, b)
    out = np.zeros(a.shape)
    inside = b < a
    aa, bb = a[inside], b[inside]
    out[inside] = (Sigma / v) ** 2 / (2 * np.pi * aa) / np.sqrt(1 - (bb / aa) ** 2) * np.exp(-aa) \
        * np.exp(np.sqrt(aa ** 2 - bb ** 2))
    return out
