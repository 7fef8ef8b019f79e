"""Device space charge (pp_cic_deposit / pp_poisson_solve / pp_efield_from_potential) against the oracle."""
import numpy as np
import pytest

from oracle import space_charge as SC
from pypatools_amd import _lib
from pypatools_amd.device import DeviceArray
from pypatools_amd.engine import ParticleState
from pypatools_amd.space_charge import SpaceChargeSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def _bunch(n, ext, seed):
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.normal(0, 0.15 * L, n) for L in ext])
    pos[:20] = rng.uniform(-2, 2, (20, 3))                 # some far outside: clipped onto the edge nodes
    return pos, rng.uniform(1e-16, 3e-16, n)


@pytest.mark.parametrize("cells", [(9, 11, 8), (16, 16, 16), (31, 7, 20)])
def test_pipeline_matches_oracle(cells):
    ext = (0.02, 0.03, 0.025)
    pos, q = _bunch(20_000, ext, 5)
    s = SpaceChargeSolver(ext, cells)
    phi, E = s.solve(pos, q)
    rho_o = SC.cic_deposit(ext, cells, pos, q)
    np.testing.assert_allclose(s.rho.to_host().reshape(cells), rho_o, rtol=1e-12, atol=1e-12 * np.abs(rho_o).max())
    phi_o = SC.solve(ext, cells, rho_o)
    np.testing.assert_allclose(phi, phi_o, rtol=1e-10, atol=1e-10 * np.abs(phi_o).max())
    # the gradient kernel on the oracle's own potential equals the oracle's gradient bit for bit
    dphi = DeviceArray.from_host(phi_o.ravel())
    out = [DeviceArray((phi_o.size,)) for _ in range(3)]
    _lib.call("pp_efield_from_potential", s.handle, dphi.ptr, *(o.ptr for o in out), None)
    for o, e in zip(out, SC.gradient(ext, cells, phi_o)):
        np.testing.assert_array_equal(o.to_host().reshape(cells), e)
    Eo = SC.gradient(ext, cells, phi)
    pts = np.column_stack([np.linspace(-0.005, 0.005, 50)] * 3)
    np.testing.assert_allclose(E(pts)[:, 0], pa_field_x(ext, cells, Eo, pts), rtol=1e-12, atol=1e-9)


def pa_field_x(ext, cells, Eo, pts):
    from oracle.field import OField
    g = dict(zip("xyz", SC.mesh_axes(ext, cells)))
    return OField.from_arrays(g, {"x": Eo[0], "y": Eo[1], "z": Eo[2]})(pts)[:, 0]


def test_solve_state_and_point_charge_symmetry():
    ext, cells = (0.04, 0.04, 0.04), (33, 33, 33)
    s = SpaceChargeSolver(ext, cells)
    h = ext[0] / cells[0]
    node = np.array([[-0.02 + 16 * h, -0.02 + 16 * h, -0.02 + 16 * h]])    # the deposit's node (16, 16, 16)
    st = ParticleState.from_host(node, np.zeros((1, 3)))
    phi, E = s.solve_state(st, 1e-12)
    assert np.unravel_index(np.argmax(phi), cells) == (16, 16, 16)
    np.testing.assert_allclose(phi, phi[::-1, :, :], rtol=1e-10)
    np.testing.assert_allclose(phi, np.transpose(phi, (1, 2, 0)), rtol=1e-10)
    Ex = s.E[0].to_host().reshape(cells)
    np.testing.assert_allclose(Ex, -Ex[::-1, :, :], rtol=1e-9, atol=1e-9 * np.abs(Ex).max())
    # far from the charge the potential approaches q / (4 pi eps0 r) minus the image of the grounded box
    assert phi[16, 16, 16] > 0 and np.all(phi >= 0)


def test_validation():
    with pytest.raises(ValueError):
        SpaceChargeSolver((0.01, 0.01, 0.01), (2, 8, 8))
    with pytest.raises(ValueError):
        SpaceChargeSolver((0.01, -0.01, 0.01), (8, 8, 8))
