"""The per-charge potential of the initial state (KMCFMM.potential, kmc_charge_potentials;
the reference has ppmd's FMM write it into group._kmc_fmm_potential at initialise,
coulomb_kmc/kmc_fmm.py:942-958):

* against the oracle's per-charge energies divided by the charge (1e-10, every
  boundary condition);
* 0.5 sum_i q_i phi_i times energy_unit equals the initialised energy (1e-10), which
  other tests pin to direct, 27-image and Ewald sums;
* free space: against the brute-force potential sum_j q_j / r_ij within the
  expansion truncation at L = 12 (1e-5 of the potential scale).
ppmd's own potential values cannot be checked here (ppmd is not available)."""

import numpy as np
import pytest

import oracle as O
from helpers import neutral_system

pytestmark = pytest.mark.gpu

EU = 1.7


@pytest.mark.parametrize("bc", ["free_space", "27", "pbc", "ff-only"])
def test_initial_potential(bc):
    from coulomb_kmc_amd import KMCFMM
    from coulomb_kmc_amd.kmc_fmm import lr_operator_reference_layout
    rng = np.random.default_rng(21)
    N, E = 300, 6.0
    pos, q = neutral_system(rng, N, E)
    k = KMCFMM(pos.copy(), q.reshape(-1, 1), E, r=4, l=12, boundary_condition=bc, max_move=1.0, energy_unit=EU,
               init_method="direct")
    k.initialise()
    phi = k.potential
    assert phi.shape == (N, 1)
    phi = phi[:, 0]
    lr = lr_operator_reference_layout(12, E) if bc in ("pbc", "ff-only") else None
    o = O.Oracle(pos, q, E, 4, 12, bc=bc, max_move=1.0, lr=lr)
    o.initialise()
    ref = o.charge_energy(np.arange(N)) / q
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(phi - ref)) <= 1e-10 * scale
    assert abs(0.5 * EU * np.dot(q, phi) - k.energy) <= 1e-10 * abs(k.energy)
    # a later call sees the current state; the initial potential stays as it was
    assert np.array_equal(k.charge_potentials(np.arange(N)), phi)
    if bc == "free_space":
        d = pos[:, None, :] - pos[None, :, :]
        r = np.sqrt((d * d).sum(-1))
        np.fill_diagonal(r, np.inf)
        direct = (q[None, :] / r).sum(axis=1)
        assert np.max(np.abs(phi - direct)) <= 1e-5 * np.max(np.abs(direct))
    k.free()
