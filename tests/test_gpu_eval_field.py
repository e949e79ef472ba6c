"""KMCFMM.eval_field (coulomb_kmc/kmc_fmm.py:544-567) against independent sums.

The reference's own checks: free space against the direct Coulomb sum at 1e-5 for
L=12 (tests/host/test_mirror_concept.py:38-100) and '27' + 'ff-only' == 'pbc' at
1e-13 (tests/host/test_split_contrib.py:110-120).  Here also '27' against the direct
27-image sum and pbc against the tinfoil Ewald potential (oracle/ewald.py)."""

import itertools

import numpy as np
import pytest

from helpers import neutral_system
from ewald import ewald_potential

pytestmark = pytest.mark.gpu


def _direct(points, pos, q, images):
    out = np.zeros(len(points))
    for o in images:
        d = points[:, None, :] - (pos[None, :, :] + np.array(o, dtype=np.float64))
        out += (q[None, :] / np.sqrt((d * d).sum(-1))).sum(1)
    return out


def _make(bc, pos, q, E, r, L, **kw):
    from coulomb_kmc_amd import KMCFMM
    k = KMCFMM(pos.copy(), q, E, r=r, l=L, boundary_condition=bc, **kw)
    k.initialise()
    return k


def test_eval_field_free_space_direct():
    rng = np.random.default_rng(1234)
    E = 4.0
    pos, q = neutral_system(rng, 50, E)
    pts = rng.uniform(-0.5 * E, 0.5 * E, (10, 3))
    k = _make("free_space", pos, q, E, 3, 12)
    phi = k.eval_field(pts)
    ref = _direct(pts, pos, q, [(0, 0, 0)])
    assert np.abs(phi - ref).max() < 1e-5
    k.free()


def test_eval_field_27_images_direct():
    rng = np.random.default_rng(8)
    E = 6.0
    pos, q = neutral_system(rng, 80, E)
    pts = rng.uniform(-0.5 * E, 0.5 * E, (12, 3))
    k = _make("27", pos, q, E, 4, 14)
    phi = k.eval_field(pts)
    ref = _direct(pts, pos, q, [tuple(E * np.array(o)) for o in itertools.product((-1, 0, 1), repeat=3)])
    assert np.abs(phi - ref).max() < 1e-6 * np.abs(ref).max()
    k.free()


def test_eval_field_pbc_ewald():
    """pbc potentials follow the cubic lattice sum with conducting boundaries (the
    long-range operator's convention): they equal the zero-mean Ewald potential minus
    the constant (2 pi / 3V) sum_j q_j |x_j|^2 (x about the box centre), which cancels
    in every energy of a neutral system and in every field."""
    rng = np.random.default_rng(21)
    E = 6.0
    pos, q = neutral_system(rng, 80, E)
    pts = rng.uniform(-0.5 * E, 0.5 * E, (12, 3))
    k = _make("pbc", pos, q, E, 4, 14)
    phi = k.eval_field(pts)
    ref = ewald_potential(pts, pos, q, E) - 2.0 * np.pi / (3.0 * E ** 3) * np.sum(q * (pos * pos).sum(1))
    assert np.abs(phi - ref).max() < 1e-5 * np.abs(ref).max()
    k.free()


def test_eval_field_split_identity():
    rng = np.random.default_rng(3)
    E = 5.0
    pos = rng.uniform(-0.5 * E, 0.5 * E, (60, 3))
    q = rng.uniform(size=60)
    q -= q.sum() / 60
    pts = rng.uniform(-0.5 * E, 0.5 * E, (60, 3))
    ks = {bc: _make(bc, pos, q, E, 3, 12) for bc in ("27", "ff-only", "pbc")}
    near, far, full = (ks[b].eval_field(pts) for b in ("27", "ff-only", "pbc"))
    assert np.abs(near + far - full).max() < 1e-13 * max(1.0, np.abs(full).max())
    for k in ks.values():
        k.free()
