"""Normwise error metrics used to report every result against a reference.

For a field of n vectors (components along axis 0, e.g. velocity (3, n) or the
gradient tensor (9, n)):
  max_rel = max_i |a_i - b_i| / max_i |b_i|        (vector 2-norms per target)
  rms_rel = sqrt(sum_i |a_i - b_i|^2 / sum_i |b_i|^2)
Normwise, because many gradient components vanish by symmetry (trace grad u = 0)
and component-wise relative error would be dominated by those zeros.
"""
import numpy as np


def rel_errors(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(a - b, axis=0)
    r = np.linalg.norm(b, axis=0)
    rmax = r.max() if r.size else 0.0
    if rmax == 0.0:
        return (0.0 if d.max(initial=0.0) == 0.0 else np.inf), 0.0
    return float(d.max() / rmax), float(np.sqrt((d ** 2).sum() / (r ** 2).sum()))


def field_errors(vel, grad, vref, gref):
    out = {}
    out["vel_max_rel"], out["vel_rms_rel"] = rel_errors(vel, vref)
    if grad is not None and gref is not None:
        out["grad_max_rel"], out["grad_rms_rel"] = rel_errors(grad, gref)
    return out


def self_term_gradient(core, sstr, srad):
    """Gradient contribution of the self pair j == i, K(0; s, s) skew(alpha_i),
    in ug order (Kernels.h:183-192 at d = 0).  K(0) per core: RM (2s^2)^-3/2,
    WL 2.5 (2s^2)^-3/2, EXP 1/(2 s^3), V2 (2 s^4)^-3/4."""
    s = np.asarray(srad, dtype=np.float64)
    k = {"rm": (2 * s * s) ** -1.5, "wl": 2.5 * (2 * s * s) ** -1.5,
         "exp": 1.0 / (2 * s ** 3), "v2": (2 * s ** 4) ** -0.75}[core]
    a = np.asarray(sstr, dtype=np.float64)
    g = np.zeros((9, s.shape[0]))
    g[1], g[2], g[3] = a[2] * k, -a[1] * k, -a[2] * k
    g[5], g[6], g[7] = a[0] * k, a[1] * k, -a[0] * k
    return g


def nerr(a, b, mask=None):
    """Entrywise normwise error of a matrix block: max |a - b| / max |b| over the
    entries where mask (default: all) is true."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is None:
        mask = np.ones(b.shape, bool)
    den = np.abs(np.where(mask, b, 0.0)).max()
    return float(np.abs(np.where(mask, a - b, 0.0)).max() / den) if den > 0 else 0.0
