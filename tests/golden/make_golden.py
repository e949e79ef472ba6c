"""
Writes the golden fixtures used by tests/test_oracle_basis.py.

The reference stores no golden files; its basis tests draw seeded inputs
(numpy RandomState) and compare against closed forms built on
scipy.special.lpmv.  This script regenerates those inputs with the same seeds
and shapes and stores the closed-form outputs:

  sph_harm_L20.npz   tests/host/test_c_sph_harm.py:45-78   (RandomState(9476213), N=20, L=20)
                     energies q * Re sum L_lm r^l Y_l^m at (x - centre), compute_phi_local form
  dot_vec_L20.npz    tests/host/test_c_sph_harm.py:129-165 (50 random expansions / points)
  multipole_L26.npz  tests/host/test_pbc_1.py:38-62        (RandomState(235243095), L=26, N=10)
  local_exp_L12.npz  tests/host/test_c_sph_harm.py:172-188 (RandomState(9476213), L=12, 10 points)

Closed form: Y_l^m = sqrt((l-|m|)!/(l+|m|)!) P_l^{|m|}(cos theta) e^{i m phi} with the
Condon-Shortley phase (scipy.special.lpmv's convention), index l*l+l+m, real block then
imaginary.  Evaluated with mpmath at 40 digits (mpmath.legenp, which carries the same
phase; checked against lpmv below), so the stored values are exact to double rounding.

Run:  python tests/golden/make_golden.py
"""

import math
import os

import mpmath as mp
import numpy as np
from scipy.special import lpmv

mp.mp.dps = 40

HERE = os.path.dirname(os.path.abspath(__file__))


def spherical(v):
    x, y, z = [mp.mpf(float(t)) for t in v]
    xy = x * x + y * y
    return mp.sqrt(xy + z * z), mp.atan2(mp.sqrt(xy), z), mp.atan2(y, x)


def Y(l, m, theta, phi):
    am = abs(m)
    h = mp.sqrt(mp.factorial(l - am) / mp.factorial(l + am))
    return h * mp.legenp(l, am, mp.cos(theta)) * mp.expj(m * phi)


def _check_phase():
    # mpmath's legenp and scipy's lpmv use the same (Condon-Shortley) phase
    for l in range(6):
        for m in range(l + 1):
            a = float(mp.legenp(l, m, mp.mpf("0.3")))
            b = float(lpmv(m, l, 0.3))
            assert abs(a - b) <= 1e-12 * max(1.0, abs(b)), (l, m, a, b)


def phi_local(L, moments, sph):
    r, th, ph = sph
    s = mp.mpf(0)
    for l in range(L):
        for m in range(-l, l + 1):
            y = Y(l, m, th, ph) * r ** l
            s += (mp.mpf(float(moments[l * l + l + m])) * mp.re(y)
                  - mp.mpf(float(moments[L * L + l * l + l + m])) * mp.im(y))
    return float(s)


def multipole(L, q, sph):
    r, th, ph = sph
    out = np.zeros(2 * L * L)
    for l in range(L):
        for m in range(-l, l + 1):
            y = Y(l, -m, th, ph) * r ** l * q
            out[l * l + l + m] = float(mp.re(y))
            out[L * L + l * l + l + m] = float(mp.im(y))
    return out


def local_exp(L, q, sph):
    r, th, ph = sph
    out = np.zeros(2 * L * L)
    for l in range(L):
        for m in range(-l, l + 1):
            y = Y(l, -m, th, ph) * q / r ** (l + 1)
            out[l * l + l + m] = float(mp.re(y))
            out[L * L + l * l + l + m] = float(mp.im(y))
    return out


def main():
    _check_phase()
    # test_c_sph_harm_1
    rng = np.random.RandomState(9476213)
    N, L = 20, 20
    ncomp = 2 * L * L
    positions = rng.uniform(low=0.0, high=1.0, size=(N, 3))
    charges = rng.uniform(size=N)
    moments = rng.uniform(low=0.0, high=1.0, size=ncomp)
    centre = np.array((0.5, 0.5, 0.5))
    energy = np.array([phi_local(L, moments, spherical(positions[i] - centre)) * float(charges[i])
                       for i in range(N)])
    np.savez(os.path.join(HERE, "sph_harm_L20.npz"), positions=positions, charges=charges, moments=moments,
             centre=centre, energy=energy, L=L)

    # test_c_local_dot_eval
    rng = np.random.RandomState(9476213)
    exps, pts, vals = [], [], []
    for _ in range(50):
        e = rng.uniform(size=ncomp)
        p = rng.uniform(size=3)
        exps.append(e)
        pts.append(p)
        vals.append(phi_local(L, e, spherical(p)))
    np.savez(os.path.join(HERE, "dot_vec_L20.npz"), expansions=np.array(exps), points=np.array(pts),
             phi=np.array(vals), L=L)

    # test_c_multipole_expansion (radius, theta, phi given directly)
    rng = np.random.RandomState(235243095)
    L2, N2 = 26, 10
    radii = rng.uniform(0.1, 10, N2)
    thetas = rng.uniform(0.001, math.pi, N2)
    phis = rng.uniform(0.001, 2.0 * math.pi, N2)
    mult = np.array([multipole(L2, 1.0, (mp.mpf(float(radii[i])), mp.mpf(float(thetas[i])), mp.mpf(float(phis[i]))))
                     for i in range(N2)])
    np.savez(os.path.join(HERE, "multipole_L26.npz"), radii=radii, thetas=thetas, phis=phis, multipole=mult, L=L2)

    # test_c_local_expansion_creation
    rng = np.random.RandomState(9476213)
    L3 = 12
    pts3 = np.array([rng.uniform(size=3) for _ in range(10)])
    lexp = np.array([local_exp(L3, 1.0, spherical(p)) for p in pts3])
    np.savez(os.path.join(HERE, "local_exp_L12.npz"), points=pts3, local=lexp, L=L3)
    print("written to", HERE)


if __name__ == "__main__":
    main()
