"""Shared test helpers: seeded systems and move streams (test infrastructure)."""

import numpy as np


def neutral_system(rng, N, E):
    """N charges +-1 (N/2 each, random order), uniform in [-E/2, E/2]^3."""
    pos = rng.uniform(-0.5 * E, 0.5 * E, (N, 3))
    q = np.ones(N)
    q[: N // 2] = -1.0
    rng.shuffle(q)
    return pos, q


def ball_moves(rng, pos, pids, k, radius, E, bc):
    """k proposals per charge, displacement uniform in a ball, wrapped (periodic) or clipped."""
    out = []
    for pid in pids:
        d = rng.normal(size=(k, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d *= radius * rng.uniform(size=(k, 1)) ** (1.0 / 3.0)
        npos = pos[pid] + d
        if bc == "free_space":
            npos = np.clip(npos, -0.5 * E, 0.5 * E)
        else:
            npos = np.where(npos < -0.5 * E, npos + E, npos)
            npos = np.where(npos >= 0.5 * E, npos - E, npos)
        out.append((int(pid), npos))
    return tuple(out)


def rel_err(a, b, floor=None):
    """The one dE error measure of the tests and bench.py (oracle.de_rel_err):
    |a - b| / max(|b|, floor), floor 1e-3 energy units unless given."""
    from oracle import de_rel_err, DE_FLOOR
    return de_rel_err(a, b, DE_FLOOR if floor is None else floor)


def metropolis_pick(dE, u, kT=1.0):
    """Index picked by u * sum(w) on the cumulative sum of w = min(1, exp(-dE/kT))."""
    dE = np.asarray(dE, dtype=np.float64)
    w = np.where(np.isfinite(dE), np.minimum(1.0, np.exp(-dE / kT)), 0.0)
    c = np.cumsum(w)
    k = int(np.searchsorted(c, u * c[-1], side="right"))
    return min(k, len(dE) - 1)


def move_stream(rng, N, P, k, radius, nsteps):
    """Seeded synthetic move stream of the benchmark configurations: per step P distinct
    moving charges, k displacements each uniform in a ball of the radius, one Metropolis
    uniform."""
    T = P * k
    pids = np.stack([rng.choice(N, size=P, replace=False) for _ in range(nsteps)]).astype(np.int64)
    d = rng.normal(size=(nsteps, T, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    d *= radius * rng.uniform(size=(nsteps, T, 1)) ** (1.0 / 3.0)
    u = rng.uniform(size=nsteps)
    return pids, d, u


def step_moves(pos, pids_s, disp_s, k, E, bc):
    """New positions of one step of a move stream (wrapped into the periodic box, or
    clipped to a free-space box), as kmc_materialize_moves forms them on the device."""
    npos = pos[np.repeat(pids_s, k)] + disp_s
    if bc == "free_space":
        return np.clip(npos, -0.5 * E, 0.5 * E)
    npos = np.where(npos < -0.5 * E, npos + E, npos)
    return np.where(npos >= 0.5 * E, npos - E, npos)
