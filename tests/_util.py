"""Shared helpers of the tests: error norms and state preparation."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def chan_errs(a, b, names):
    return {n: rel_l2(a[..., i], b[..., i]) for i, n in enumerate(names)}


def theta_pert(V, BS, nz):
    """theta minus the base-state theta0 of the level (the top texel row uses level nz-1)."""
    th0 = np.asarray(BS, np.float64)[0:8 * nz:8][:nz]
    th0 = np.concatenate([th0, th0[-1:]])
    return np.asarray(V, np.float64)[..., 3] - th0[:, None, None]
