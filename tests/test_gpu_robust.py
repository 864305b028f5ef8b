"""A state whose positions are not finite comes back as an error status, never as a fault.

NaN compares false, so a face box with a NaN corner overlaps every other box: once positions
stop being finite the candidate-pair count of a traversal grows as F^2 (1.4e11 at 513^2, beyond
the 32-bit offsets of the pair list).  The traversal caps it: more than 65,536 candidates for one
face, or more than 2^28 in one pass, end the stage with CS_E_CAPACITY before any pair is written.
A PCG whose residual turns non-finite stops at once (CS_E_NOT_CONVERGED) instead of running to
its iteration cap.
"""
import numpy as np
import pytest

from clothsimulator_amd import CS_MESH_CLOTH, ClothError, Simulator, scenes
from clothsimulator_amd._cabi import CS_E_CAPACITY, CS_E_NOT_CONVERGED, CS_OK

pytestmark = pytest.mark.gpu


def _nan_state(sim, every=True):
    x, x0, v = sim.state()
    x = x.copy()
    if every:
        x[:] = np.nan
    else:
        x[len(x) // 2] = np.nan
    sim.set_state(CS_MESH_CLOTH, 0, x, x0, v)


def test_one_nan_node_stops_the_solve_and_reports_it():
    sim = Simulator(scenes.crumple(33))
    _nan_state(sim, every=False)
    s, st = sim.step()
    print("one NaN node: status %d, pcg %d, converged %d" % (s, st.pcg_iters, st.pcg_converged))
    assert s == CS_E_NOT_CONVERGED and st.pcg_converged == 0
    assert st.pcg_iters < 50  # stopped at the breakdown, not at the 4000-iteration cap


@pytest.mark.parametrize("n", [110, 190, 513])
def test_nonfinite_positions_end_the_step_with_capacity(n):
    """110^2: 2.7e8 pairs in the pass (the 2^28 cap); 190^2 and 513^2: a face's candidates pass
    65,536 first (at 513^2 the uncapped count would wrap the 32-bit pair offsets)."""
    sim = Simulator(scenes.crumple(n))
    _nan_state(sim)
    with pytest.raises(ClothError) as e:
        sim.physics_step()
    print("crumple%d all NaN: %s" % (n, e.value), flush=True)
    assert e.value.status == CS_E_CAPACITY and "finite" in str(e.value)
    # the context is still usable once the state is finite again
    spec = scenes.crumple(n)
    m = spec.cloths[0].mesh
    sim.set_state(CS_MESH_CLOTH, 0, m.x, m.x, np.zeros_like(m.x))
    s, st = sim.step()
    assert s == CS_OK
