"""
Periodic far-field operator for the 'pbc' and 'ff-only' boundary conditions.

The reference takes this operator from ppmd's ``LongRangeMTL``
(coulomb_kmc/kmc_full_long_range.py:40, used at :186-188, :464, :882-884): a real
L^2 x L^2 matrix R applied identically to the real and imaginary halves of the
level-0 multipole M (``linop_csr_both``, :641-692) plus a three-number dipole
correction (``apply_dipole_correction_split``, :516-560).  ppmd is not part of the
reference tree, so this module builds its own operator of the same shape and meaning:

    local expansion at the origin due to every periodic image of the primary box
    outside the 3x3x3 block of boxes around it,

so that  E_pbc = E_'27' + 1/2 * (R M + dipole(M)) . M  equals the Ewald energy of the
neutral periodic system with conducting ('tinfoil') boundary conditions, which is the
convention the reference is tested against (tests/host/test_common.py:106-183,
tests/host/test_pbc_1.py:91-144).

Method (independent of ppmd's):

* M2L in Racah-scaled solid harmonics:  for a multipole Q about centre c,
  local coefficients about the origin are  L~_jk = sum_lm (-1)^l Q_lm I~_{l+j,m+k}(c).
  Summed over the far lattice this needs only the lattice sums
  S_p^q = sum_{|n|_inf >= 2} I~_p^q(n E).
* The lattice sums are evaluated in expanding-cube order by renormalisation
  (the far set splits into the ring 2 <= |n|_inf <= 4 and a copy of itself at
  three times the scale; the 27 boxes of a super-cell are folded with the
  irregular translation theorem).  Odd p vanish by inversion symmetry, q != 0 mod 4
  by the fourfold axis, p <= 2 are the conditionally convergent sums that vanish on
  every cubic shell.
* Expanding cubes give the vacuum-surrounded (spherical-shape) result; the
  conversion to tinfoil is the dipole term -2 pi |D|^2 / (3 V), written in the
  reference's (SX, SY, SZ) form.
"""

import functools
import math
import numpy as np

from .harmonics import regular_columns, irregular_columns, half_index, half_size


# ----------------------------------------------------------------------------
# Racah-scaled harmonics over -l <= m <= l, array index [..., l, m + lmax]
# ----------------------------------------------------------------------------

def _lognorm_table(lmax):
    """log N_lm = 0.5*(log (l+m)! + log (l-m)!) for |m| <= l <= lmax."""
    lf = np.array([math.lgamma(k + 1.0) for k in range(2 * lmax + 2)])
    t = np.full((lmax + 1, 2 * lmax + 1), -np.inf)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            t[l, m + lmax] = 0.5 * (lf[l + m] + lf[l - m])
    return t


def _eps(m):
    return -1.0 if (m < 0 and (m % 2) == 1) else 1.0


def racah_regular(lmax, pts):
    """R~_l^m(x) = eps_m R_l^m(x) / N_lm for all |m| <= l <= lmax."""
    cols = regular_columns(lmax, pts)
    lnN = _lognorm_table(lmax)
    out = np.zeros(cols.shape[:-2] + (lmax + 1, 2 * lmax + 1), dtype=np.complex128)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            v = cols[..., l, abs(m)]
            if m < 0:
                v = np.conj(v)
            out[..., l, m + lmax] = _eps(m) * v * math.exp(-lnN[l, m + lmax])
    return out


def racah_irregular(lmax, pts):
    """I~_l^m(x) = eps_m I_l^m(x) * N_lm for all |m| <= l <= lmax."""
    cols = irregular_columns(lmax, pts)
    lnN = _lognorm_table(lmax)
    out = np.zeros(cols.shape[:-2] + (lmax + 1, 2 * lmax + 1), dtype=np.complex128)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            v = cols[..., l, abs(m)]
            if m < 0:
                v = np.conj(v)
            out[..., l, m + lmax] = _eps(m) * v * math.exp(lnN[l, m + lmax])
    return out


# ----------------------------------------------------------------------------
# lattice sums
# ----------------------------------------------------------------------------

def _ring_vectors(lo, hi):
    r = np.arange(-hi, hi + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    ninf = np.abs(g).max(axis=1)
    return g[(ninf >= lo) & (ninf <= hi)].astype(np.float64)


@functools.lru_cache(maxsize=8)
def far_lattice_sums(pmax, extra=48):
    """S_p^q = sum over n in Z^3 with |n|_inf >= 2 of I~_p^q(n) (unit box), cube order.

    Returns a real array [p, q + pmax] for 0 <= p <= pmax.  Entries that vanish by the
    cubic point group (odd p, q != 0 mod 4, p <= 2) are exactly zero.
    """
    ptop = pmax + extra
    ring = racah_irregular(ptop, _ring_vectors(2, 4)).sum(axis=0)
    cube = _ring_vectors(0, 1)
    creg = racah_regular(extra, cube).conj().sum(axis=0)          # C_jk = sum_o R~*_jk(o)

    S = np.zeros((ptop + 1, 2 * ptop + 1), dtype=np.complex128)
    for p in range(ptop, 2, -1):
        if p % 2 == 1:
            continue
        denom = 1.0 - 27.0 * 3.0 ** (-(p + 1))
        for q in range(-p, p + 1):
            if q % 4 != 0:
                continue
            acc = ring[p, q + ptop]
            for j in range(4, min(extra, ptop - p) + 1, 2):
                scale = 3.0 ** (-(p + j + 1))
                for k in range(-j, j + 1):
                    if k % 4 != 0 or abs(q + k) > p + j:
                        continue
                    c = creg[j, k + extra]
                    if c != 0.0:
                        acc += c * scale * S[p + j, q + k + ptop]
            S[p, q + ptop] = acc / denom
    out = np.zeros((pmax + 1, 2 * pmax + 1))
    for p in range(pmax + 1):
        for q in range(-p, p + 1):
            out[p, q + pmax] = S[p, q + ptop].real
    out.setflags(write=False)   # cached: shared by every caller
    return out


# ----------------------------------------------------------------------------
# the operator in the reference's layouts
# ----------------------------------------------------------------------------

def far_field_matrix(L, extent):
    """R as a dense real L^2 x L^2 matrix in the reference's (l, m) index l*l+l+m.

    Applied to the real and to the imaginary half of a multipole expansion M about
    the origin it gives the local expansion at the origin of all images beyond the
    3x3x3 block (coulomb_kmc/kmc_full_long_range.py:641-692 applies it this way).
    """
    pmax = 2 * L - 2
    S = far_lattice_sums(pmax)
    lnN = _lognorm_table(L)
    R = np.zeros((L * L, L * L))
    for j in range(L):
        for k in range(-j, j + 1):
            row = j * j + j + k
            for l in range(L):
                p = j + l
                for m in range(-l, l + 1):
                    q = m - k
                    if abs(q) > p:
                        continue
                    s = S[p, q + pmax]
                    if s == 0.0:
                        continue
                    val = ((-1.0) ** l) * _eps(-k) * _eps(m) * s
                    val *= math.exp(-lnN[j, k + L] - lnN[l, m + L])
                    val *= extent ** (-(p + 1))
                    R[row, l * l + l + m] = val
    return R


def dipole_correction(extent):
    """(SX, SY, SZ) of the reference's dipole term (kmc_full_long_range.py:516-560).

    The term 0.5*(SX*M_re11*(E_re11+E_re1-1) - SY*M_im11*E_im11 + SY*M_im11*E_im1-1
    + SZ*M_re10*E_re10) equals -2 pi |D|^2 / (3 V), converting the expanding-cube
    lattice sum to conducting boundary conditions.
    """
    v = extent ** 3
    c = 4.0 * math.pi / (3.0 * v)
    return np.array([-c, c, -c])


def full_operator(L, extent):
    """The 2L^2 x 2L^2 real operator F with  E_lr(M) = 0.5 * M . (F M)  (reference layout)."""
    n = L * L
    R = far_field_matrix(L, extent)
    F = np.zeros((2 * n, 2 * n))
    F[:n, :n] = R
    F[n:, n:] = R
    if L > 1:
        sx, sy, sz = dipole_correction(extent)
        re11, re10, re1n1 = 1 + 1 + 1, 1 + 1 + 0, 1 + 1 - 1
        im11, im1n1 = n + re11, n + re1n1
        F[re11, re11] += sx
        F[re1n1, re11] += sx
        F[im11, im11] -= sy
        F[im1n1, im11] += sy
        F[re10, re10] += sz
    return F


def half_operator(L, extent):
    """Symmetric operator G on the packed (l, m>=0) (re, im) vector with
    0.5 * h . (G h) == 0.5 * M . (F M) for every real-charge expansion M.

    Packed vector layout: index 2*half_index(l, m) + c, c = 0 real, 1 imaginary.
    """
    L2 = L * L
    nh = half_size(L)
    F = full_operator(L, extent)
    P = np.zeros((2 * L2, 2 * nh))
    for l in range(L):
        for m in range(-l, l + 1):
            h = half_index(l, abs(m))
            P[l * l + l + m, 2 * h] = 1.0
            P[L2 + l * l + l + m, 2 * h + 1] = 1.0 if m >= 0 else -1.0
    G = P.T @ F @ P
    return 0.5 * (G + G.T)


def csr(mat, tol=0.0):
    """(data, indptr, indices) of the rows of a dense matrix, keeping |a| > tol."""
    data, indices, indptr = [], [], [0]
    for row in mat:
        nz = np.nonzero(np.abs(row) > tol)[0]
        indices.extend(nz.tolist())
        data.extend(row[nz].tolist())
        indptr.append(len(indices))
    return (np.array(data, dtype=np.float64), np.array(indptr, dtype=np.int64),
            np.array(indices, dtype=np.int64))
