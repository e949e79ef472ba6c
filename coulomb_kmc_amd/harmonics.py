"""
Host-side (numpy) solid harmonics in the coefficient convention of the
reference's expansion tools (coulomb_kmc/kmc_expansion_tools.py:17-491).

Convention (the one the reference's local_eval / create_local_exp / multipole_exp /
local_dot_vec follow; pinned in coulomb_kmc/kmc_expansion_tools.py:384-419 and
tests/host/test_c_sph_harm.py:129-165 against scipy.special.lpmv):

    Y_l^m(theta, phi) = sqrt((l-|m|)!/(l+|m|)!) * P_l^{|m|}(cos theta) * exp(i m phi)

with P_l^m carrying the Condon-Shortley phase (scipy.special.lpmv), so that
Y_l^{-m} = conj(Y_l^m).  Regular and irregular solid harmonics are

    R_l^m(x) = r^l Y_l^m(x^),      I_l^m(x) = Y_l^m(x^) / r^(l+1).

Both are evaluated without trigonometry through the column recurrence (m >= 0)

    b_m^m = s_m,   b_l^m = alpha_lm z b_{l-1}^m - beta_lm r^2 b_{l-2}^m,
    R_l^m = b_l^m (x + i y)^m

    c_m^m = s_m / r^(2m+1),   c_l^m = (alpha_lm z c_{l-1}^m - beta_lm c_{l-2}^m) / r^2,
    I_l^m = c_l^m (x + i y)^m

    alpha_lm = (2l-1)/sqrt(l^2-m^2),  beta_lm = sqrt(((l-1)^2-m^2)/(l^2-m^2)),
    s_m = (-1)^m (2m-1)!!/sqrt((2m)!)

which is the Legendre three-term recurrence with the normalisation folded in.  The
GPU kernels use the same recurrence and the same coefficient tables (built by
``recurrence_tables``).

This module is used at set-up time only (the PBC lattice operator); nothing on the
propose/accept path calls it.
"""

import math
import numpy as np


def recurrence_tables(lmax):
    """alpha[l, m], beta[l, m] and s[m] for 0 <= m <= l <= lmax (zeros elsewhere)."""
    alpha = np.zeros((lmax + 1, lmax + 1))
    beta = np.zeros((lmax + 1, lmax + 1))
    s = np.zeros(lmax + 1)
    s[0] = 1.0
    for m in range(1, lmax + 1):
        s[m] = -s[m - 1] * math.sqrt((2.0 * m - 1.0) / (2.0 * m))
    for m in range(lmax + 1):
        for l in range(m + 1, lmax + 1):
            d = float(l * l - m * m)
            alpha[l, m] = (2.0 * l - 1.0) / math.sqrt(d)
            beta[l, m] = math.sqrt(max(0.0, float((l - 1) * (l - 1) - m * m)) / d)
    return alpha, beta, s


def regular_columns(lmax, pts):
    """R_l^m(x) for 0 <= m <= l <= lmax at points pts (..., 3).

    Returns a complex array of shape (..., lmax+1, lmax+1) indexed [l, m], m >= 0;
    entries with m > l are zero.  R_l^{-m} = conj(R_l^m).
    """
    pts = np.asarray(pts, dtype=np.float64)
    x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
    r2 = x * x + y * y + z * z
    alpha, beta, s = recurrence_tables(lmax)
    out = np.zeros(pts.shape[:-1] + (lmax + 1, lmax + 1), dtype=np.complex128)
    w = x + 1j * y
    wm = np.ones_like(w)
    for m in range(lmax + 1):
        b2 = np.zeros_like(x)
        b1 = np.full_like(x, s[m])
        out[..., m, m] = b1 * wm
        for l in range(m + 1, lmax + 1):
            b = alpha[l, m] * z * b1 - beta[l, m] * r2 * b2
            out[..., l, m] = b * wm
            b2, b1 = b1, b
        wm = wm * w
    return out


def irregular_columns(lmax, pts):
    """I_l^m(x) for 0 <= m <= l <= lmax at points pts (..., 3); layout as regular_columns."""
    pts = np.asarray(pts, dtype=np.float64)
    x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
    r2 = x * x + y * y + z * z
    ir2 = 1.0 / r2
    ir = np.sqrt(ir2)
    alpha, beta, s = recurrence_tables(lmax)
    out = np.zeros(pts.shape[:-1] + (lmax + 1, lmax + 1), dtype=np.complex128)
    u = (x + 1j * y) * ir2
    um = ir.astype(np.complex128)
    for m in range(lmax + 1):
        c2 = np.zeros_like(x)
        c1 = np.full_like(x, s[m])
        out[..., m, m] = c1 * um
        for l in range(m + 1, lmax + 1):
            c = (alpha[l, m] * z * c1 - beta[l, m] * c2) * ir2
            out[..., l, m] = c * um
            c2, c1 = c1, c
        um = um * u
    return out


def full_from_columns(cols, L):
    """Expand [l, m>=0] columns to the reference's (re, im) layout over -l <= m <= l.

    Index of (l, m) is l*l + l + m (coulomb_kmc/kmc_expansion_tools.py:42-43); the
    result has shape (..., 2*L*L): real parts first, then imaginary parts.
    Uses X_l^{-m} = conj(X_l^m).
    """
    shp = cols.shape[:-2]
    out = np.zeros(shp + (2 * L * L,))
    for l in range(L):
        for m in range(-l, l + 1):
            v = cols[..., l, abs(m)]
            if m < 0:
                v = np.conj(v)
            out[..., l * l + l + m] = v.real
            out[..., L * L + l * l + l + m] = v.imag
    return out


def half_index(l, m):
    """Position of (l, m >= 0) in the packed half layout used on the GPU."""
    return l * (l + 1) // 2 + m


def half_size(L):
    return L * (L + 1) // 2


def expansion_full_to_half(full, L):
    """Reference (re | im) layout -> packed (l, m>=0) complex pairs, shape (..., L(L+1)/2, 2).

    Only the m >= 0 entries are read; for expansions built from real charges the
    m < 0 entries are their complex conjugates.
    """
    full = np.asarray(full, dtype=np.float64)
    nh = half_size(L)
    out = np.zeros(full.shape[:-1] + (nh, 2))
    for l in range(L):
        for m in range(l + 1):
            out[..., half_index(l, m), 0] = full[..., l * l + l + m]
            out[..., half_index(l, m), 1] = full[..., L * L + l * l + l + m]
    return out


def expansion_half_to_full(half, L):
    """Packed (l, m>=0) complex pairs -> reference (re | im) layout (conjugate fill for m<0)."""
    half = np.asarray(half, dtype=np.float64)
    out = np.zeros(half.shape[:-2] + (2 * L * L,))
    for l in range(L):
        for m in range(-l, l + 1):
            h = half[..., half_index(l, abs(m)), :]
            out[..., l * l + l + m] = h[..., 0]
            out[..., L * L + l * l + l + m] = h[..., 1] if m >= 0 else -h[..., 1]
    return out
