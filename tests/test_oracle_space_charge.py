"""The space-charge oracle (poisson_amg.py restated for a box without electrodes),
pinned by the discrete eigenfunction solution, charge conservation and exact gradients."""
import numpy as np

from oracle import space_charge as SC

EXT, CELLS = (0.02, 0.03, 0.025), (9, 11, 8)


def test_discrete_eigenfunction_is_solved_exactly():
    nx, ny, nz = CELLS
    h = [L / n for L, n in zip(EXT, CELLS)]
    kx, ky, kz = 2, 3, 1
    sx = np.sin(np.pi * kx * np.arange(1, nx + 1) / (nx + 1))
    sy = np.sin(np.pi * ky * np.arange(1, ny + 1) / (ny + 1))
    sz = np.sin(np.pi * kz * np.arange(1, nz + 1) / (nz + 1))
    mode = sx[:, None, None] * sy[None, :, None] * sz[None, None, :]
    lam = sum(4.0 / hh ** 2 * np.sin(np.pi * k / (2 * (n + 1))) ** 2 for hh, k, n in zip(h, (kx, ky, kz), CELLS))
    rho = 1e-9 * mode
    phi = SC.solve(EXT, CELLS, rho)
    np.testing.assert_allclose(phi, rho / SC.EPS0 / lam, rtol=1e-10, atol=1e-12 * np.abs(phi).max())
    # and the operator is the 7-point stencil: A phi reproduces rho / eps0
    A = SC.operator(EXT, CELLS)
    np.testing.assert_allclose(A @ phi.ravel(), rho.ravel() / SC.EPS0, rtol=1e-9, atol=1e-9 * np.abs(rho).max() / SC.EPS0)


def test_deposit_conserves_charge_and_hits_nodes():
    rng = np.random.default_rng(3)
    h = [L / n for L, n in zip(EXT, CELLS)]
    pos = np.column_stack([rng.uniform(-L / 2 + 0.01 * L, L / 2 - h_ - 0.01 * L, 500) for L, h_ in zip(EXT, h)])
    q = rng.uniform(1e-15, 2e-15, 500)
    rho = SC.cic_deposit(EXT, CELLS, pos, q)
    np.testing.assert_allclose(rho.sum() * h[0] * h[1] * h[2], q.sum(), rtol=1e-12)
    # the deposit's node i sits at -L/2 + i h (half a cell from the mesh's cell centres)
    node = np.array([[-EXT[0] / 2 + 3 * h[0], -EXT[1] / 2 + 5 * h[1], -EXT[2] / 2 + 2 * h[2]]])
    r1 = SC.cic_deposit(EXT, CELLS, node, np.array([1e-12]))
    assert np.isclose(r1[3, 5, 2] * h[0] * h[1] * h[2], 1e-12, rtol=1e-14, atol=0) and np.count_nonzero(r1) == 1
    # beyond the upper edge everything lands on node n - 2 (clip to n - 1, f = 0 kept)
    far = np.array([[1.0, 1.0, 1.0]])
    r2 = SC.cic_deposit(EXT, CELLS, far, np.array([1.0]))
    assert np.count_nonzero(r2) == 1 and r2[-2, -2, -2] > 0


def test_gradient_of_linear_potential():
    x, y, z = SC.mesh_axes(EXT, CELLS)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    phi = 3.0 * X - 2.0 * Y + 0.5 * Z
    E = SC.gradient(EXT, CELLS, phi)
    c = (slice(1, -1),) * 3
    np.testing.assert_allclose(E[0][c], -3.0, rtol=1e-12)
    np.testing.assert_allclose(E[1][c], 2.0, rtol=1e-12)
    np.testing.assert_allclose(E[2][c], -0.5, rtol=1e-12)
    assert np.all(E[0][0] == 0) and np.all(E[2][:, :, -1] == 0)
