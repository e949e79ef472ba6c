"""
Oracle (test infrastructure only): independent Coulomb energies used to pin the
periodic far-field operator and the free-space / '27' energies.

Nothing on the product path imports this module.

* ``free_space_energy``  : sum_{i<j} q_i q_j / r_ij  (the reference's FreeSpaceDirect
  formula, restated in tests/host/test_common.py:36-43).
* ``nearest27_energy``   : primary box plus its 26 images in vacuum
  (NearestDirect, restated in tests/host/test_common.py:76-90).
* ``ewald_energy``       : Ewald sum of the neutral periodic system with conducting
  (tinfoil) boundary conditions - the convention of the reference's
  EwaldOrthoganalHalf checker (tests/host/test_common.py:106-183).  Textbook split
  (Ewald 1921; e.g. Frenkel & Smit, ch. 12):
      U = 1/2 sum_{i,j,n}' q_i q_j erfc(a r)/r + (2 pi/V) sum_{k!=0} e^{-k^2/4a^2}/k^2 |S(k)|^2
          - a/sqrt(pi) sum q_i^2
"""

import itertools
import math

import numpy as np
from scipy.special import erfc


def free_space_energy(pos, q):
    pos = np.asarray(pos, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64).ravel()
    e = 0.0
    for i in range(len(q)):
        d = np.linalg.norm(pos[i + 1:] - pos[i], axis=1)
        e += q[i] * np.dot(q[i + 1:], 1.0 / d)
    return e


def nearest27_energy(pos, q, extent):
    pos = np.asarray(pos, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64).ravel()
    e = free_space_energy(pos, q)
    for o in itertools.product((-1, 0, 1), repeat=3):
        if o == (0, 0, 0):
            continue
        shift = extent * np.array(o, dtype=np.float64)
        for i in range(len(q)):
            d = np.linalg.norm(pos + shift - pos[i], axis=1)
            e += 0.5 * q[i] * np.dot(q, 1.0 / d)
    return e


def ewald_energy(pos, q, extent, tol=1e-15):
    """Tinfoil Ewald energy of a neutral cubic periodic system (all O(N^2) sums)."""
    pos = np.asarray(pos, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64).ravel()
    E = float(extent)
    V = E ** 3
    alpha = 5.0 / E
    # real space: erfc(alpha r) < tol beyond r_c
    rc = 1.0
    while erfc(alpha * rc) / rc > tol:
        rc *= 1.1
    nmax = int(math.ceil(rc / E)) + 1
    e_real = 0.0
    diffs = pos[:, None, :] - pos[None, :, :]
    for n in itertools.product(range(-nmax, nmax + 1), repeat=3):
        d = diffs + E * np.array(n, dtype=np.float64)
        r = np.sqrt((d * d).sum(axis=-1))
        if n == (0, 0, 0):
            np.fill_diagonal(r, np.inf)
        mask = r < rc
        e_real += 0.5 * np.sum(np.outer(q, q)[mask] * erfc(alpha * r[mask]) / r[mask])
    # reciprocal space: exp(-k^2/4a^2) < tol beyond k_c
    kc = 2.0 * alpha * math.sqrt(-math.log(tol))
    kmax = int(math.ceil(kc * E / (2.0 * math.pi)))
    e_rec = 0.0
    for kv in itertools.product(range(-kmax, kmax + 1), repeat=3):
        if kv == (0, 0, 0):
            continue
        k = (2.0 * math.pi / E) * np.array(kv, dtype=np.float64)
        k2 = float(k @ k)
        if k2 > kc * kc:
            continue
        s = np.sum(q * np.exp(1j * (pos @ k)))
        e_rec += math.exp(-k2 / (4.0 * alpha * alpha)) / k2 * abs(s) ** 2
    e_rec *= 2.0 * math.pi / V
    e_self = -alpha / math.sqrt(math.pi) * float(np.sum(q * q))
    return e_real + e_rec + e_self


def ewald_potential(points, pos, q, extent, tol=1e-15):
    """Tinfoil Ewald potential phi(x) = sum_{j, n} q_j / |x - x_j + n E| of a neutral cubic
    periodic system at points that are not charge positions (same split as ewald_energy;
    the uniform k = 0 term vanishes for a neutral system)."""
    points = np.atleast_2d(np.asarray(points, dtype=np.float64))
    pos = np.asarray(pos, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64).ravel()
    E = float(extent)
    V = E ** 3
    alpha = 5.0 / E
    rc = 1.0
    while erfc(alpha * rc) / rc > tol:
        rc *= 1.1
    nmax = int(math.ceil(rc / E)) + 1
    phi = np.zeros(len(points))
    diffs = points[:, None, :] - pos[None, :, :]
    for n in itertools.product(range(-nmax, nmax + 1), repeat=3):
        d = diffs + E * np.array(n, dtype=np.float64)
        r = np.sqrt((d * d).sum(axis=-1))
        mask = r < rc
        phi += np.sum(np.where(mask, q[None, :] * erfc(alpha * r) / np.where(mask, r, 1.0), 0.0), axis=1)
    kc = 2.0 * alpha * math.sqrt(-math.log(tol))
    kmax = int(math.ceil(kc * E / (2.0 * math.pi)))
    for kv in itertools.product(range(-kmax, kmax + 1), repeat=3):
        if kv == (0, 0, 0):
            continue
        k = (2.0 * math.pi / E) * np.array(kv, dtype=np.float64)
        k2 = float(k @ k)
        if k2 > kc * kc:
            continue
        s = np.sum(q * np.exp(-1j * (pos @ k)))
        phi += (4.0 * math.pi / V) * math.exp(-k2 / (4.0 * alpha * alpha)) / k2 * np.real(np.exp(1j * (points @ k)) * s)
    return phi
