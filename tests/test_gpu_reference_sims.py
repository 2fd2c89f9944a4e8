"""
The reference's own simulation test, every row, on the device:
tests/mesh/test_sims.py:30-82 of the reference, parametrized over
{length gauge + AlternatingDirectionImplicit, length gauge + SplitInteractionOperator,
velocity gauge + SplitInteractionOperator} x the low-n hydrogen bound states (1, 0), (2, 0),
(2, 1) (tests/mesh/conftest.py:24-26). With no potential the numeric eigenstate is stationary:
the norm and every state overlap are the same at the first and last data times to 1e-14.
"""

import numpy as np
import pytest

from ionization_amd import mesh, states
from ionization_amd import units as u

pytestmark = pytest.mark.gpu

LOW_N_HYDROGEN_BOUND_STATES = [states.HydrogenBoundState(n, l) for n in range(3) for l in range(n)]

ROWS = {
    "LEN-ADI": (mesh.SphericalHarmonicLengthGaugeOperators, mesh.AlternatingDirectionImplicit),
    "LEN-SIO": (mesh.SphericalHarmonicLengthGaugeOperators, mesh.SplitInteractionOperator),
    "VEL-SIO": (mesh.SphericalHarmonicVelocityGaugeOperators, mesh.SplitInteractionOperator),
}


@pytest.mark.parametrize("initial_state", LOW_N_HYDROGEN_BOUND_STATES, ids=str)
@pytest.mark.parametrize("row", sorted(ROWS))
def test_with_no_potential_final_state_is_initial_state_with_numeric_eigenstates(initial_state, row):
    operators, evolution_method = ROWS[row]
    sim = mesh.SphericalHarmonicSpecification(
        "test",
        initial_state=initial_state,
        operators=operators(),
        evolution_method=evolution_method(),
        time_initial=0,
        time_final=100 * u.asec,
        time_step=1 * u.asec,
        r_bound=50 * u.bohr_radius,
        r_points=250,
        l_bound=30,
        use_numeric_eigenstates=True,
        numeric_eigenstate_max_energy=10 * u.eV,
        numeric_eigenstate_max_angular_momentum=10,
    ).to_sim()
    assert len(LOW_N_HYDROGEN_BOUND_STATES) == 3

    sim.run()

    initial_norm = sim.data.norm[0]
    final_norm = sim.data.norm[-1]
    initial_state_overlaps = np.array([v[0] for v in sim.data.state_overlaps.values()])
    final_state_overlaps = np.array([v[-1] for v in sim.data.state_overlaps.values()])

    assert np.allclose(initial_norm, final_norm, atol=1e-14)
    assert np.allclose(initial_state_overlaps, final_state_overlaps, atol=1e-14)
    # the initial state is one of the numeric test states and stays fully populated
    assert max(final_state_overlaps) == pytest.approx(1.0, abs=1e-12)
