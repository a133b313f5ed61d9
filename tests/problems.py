"""Seeded point sets and densities shared by the tests (numpy generators only)."""
import numpy as np

COEFS = {111: [1.0], 121: [1.0], 311: [1.3], 312: [1.3], 321: [1.0], 322: [1.3], 511: [1.0, 0.3], 521: [1.2, 0.3]}


def uniform_cube(n, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3))


def clusters(n, seed, k=6, lo=0.01, hi=0.1):
    """Gaussian clusters inside [0,1)^3 with log-uniform widths; samples outside are redrawn."""
    rng = np.random.default_rng(seed)
    ctr = rng.uniform(0.1, 0.9, (k, 3))
    sig = np.exp(rng.uniform(np.log(lo), np.log(hi), k))
    out = np.empty((0, 3))
    while out.shape[0] < n:
        lab = rng.integers(0, k, n)
        p = ctr[lab] + sig[lab, None] * rng.standard_normal((n, 3))
        p = p[np.all((p >= 0.0) & (p < 1.0), axis=1)]
        out = np.concatenate([out, p])
    return out[:n]


def sphere(n, seed, radius=1.0):
    """Points on a sphere (uniform in cos(theta) and phi) and outward unit normals."""
    rng = np.random.default_rng(seed)
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    nor = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    return radius * nor, nor


def densities(n, dof, nrhs, seed):
    rng = np.random.default_rng(seed + 7)
    return rng.uniform(-1, 1, (nrhs, n * dof))


def relerr(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
