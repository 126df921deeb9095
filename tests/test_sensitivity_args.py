"""Argument checks of z4c_amd.sensitivity.hamiltonian_sensitivity that run
before any device call: the adjoint window must lie within the evolution,
and the outflow fill is only part of a step that runs the exchange (the
step VJP transposes the fill after the update's transpose)."""
import pytest

from z4c_amd.sensitivity import hamiltonian_sensitivity


@pytest.mark.parametrize("nsteps,window", [(4, 0), (4, 5), (0, 0)])
def test_window_must_lie_within_the_evolution(nsteps, window):
    with pytest.raises(ValueError):
        hamiltonian_sensitivity(None, None, 0.1, nsteps, window)


def test_outflow_fill_needs_the_exchange():
    with pytest.raises(ValueError):
        hamiltonian_sensitivity(None, None, 0.1, 4, 2, physical=True)
