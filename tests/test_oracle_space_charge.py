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


# --------------------------------------------------------------------------- electrode-bounded domains

from pypatools_amd.electrodes import box_mesh, sphere_mesh      # noqa: E402  (numpy only)

EXT, CELLS = (0.1, 0.08, 0.06), (10, 8, 6)


def _nodes(extent, cells):
    ax = SC.mesh_axes(extent, cells)
    return [g.ravel() for g in np.meshgrid(*ax, indexing="ij")]


def test_axis_rays_half_space_box():
    """A box electrode filling x < c across the domain: conductor exactly where x < c; the distances along x
    to the face are x - c; crossing counts are those of two faces per line."""
    c = 0.0123
    v, t = box_mesh([-1.0, -1.0, -1.0], [c, 1.0, 1.0])
    dist, counts = SC.axis_ray_queries(v, t, EXT, CELLS)
    x, y, z = _nodes(EXT, CELLS)
    inside = x < c
    assert np.array_equal(SC.classify(EXT, CELLS, dist, counts) == SC.CONDUCTOR, inside)
    np.testing.assert_array_equal(counts[~inside], np.tile([0, 2, 0, 0, 0, 0], ((~inside).sum(), 1)))
    np.testing.assert_array_equal(counts[inside], np.tile([1, 1, 1, 1, 1, 1], (inside.sum(), 1)))
    np.testing.assert_allclose(dist[~inside, 1], x[~inside] - c, rtol=1e-13)
    np.testing.assert_allclose(dist[inside, 0], c - x[inside], rtol=1e-13)
    assert np.all(np.isinf(dist[~inside][:, [0, 2, 3, 4, 5]]))


def test_axis_rays_through_face_diagonals_count_once():
    """Cube centred on a symmetric grid: the lines y = z (and x = y, z = x) run exactly through the face
    diagonals shared by two triangles.  Each face is still crossed once, so the classification is exact."""
    ext, cells = (0.08, 0.08, 0.08), (8, 8, 8)
    s = 0.023
    v, t = box_mesh([-s, -s, -s], [s, s, s])
    dist, counts = SC.axis_ray_queries(v, t, ext, cells)
    x, y, z = _nodes(ext, cells)
    inside = (np.abs(x) < s) & (np.abs(y) < s) & (np.abs(z) < s)
    on_diag = (y == z) & (np.abs(y) < s)
    assert on_diag.sum() > 0
    assert np.array_equal(SC.classify(ext, cells, dist, counts) == SC.CONDUCTOR, inside)
    for d in range(6):
        assert set(np.unique(counts[:, d])) <= {0, 1, 2}
    np.testing.assert_array_equal(counts[inside], 1)
    line_hits = (np.abs(y) < s) & (np.abs(z) < s)            # x-lines through the cube
    np.testing.assert_array_equal(counts[line_hits, 0] + counts[line_hits, 1], 2)


def test_axis_rays_icosphere_against_plane_geometry():
    """Convex icosphere: conductor = inside every face plane; ray distances equal the ray-plane distance of
    the nearest face whose triangle contains the hit point (computed here by plane and barycentric tests)."""
    ext, cells = (0.1, 0.1, 0.1), (14, 14, 14)
    v, t = sphere_mesh([0.003, -0.002, 0.001], 0.031, 2)
    dist, counts = SC.axis_ray_queries(v, t, ext, cells)
    P = np.column_stack(_nodes(ext, cells))
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    nrm = np.cross(b - a, c - a)
    off = np.einsum("ij,ij->i", nrm, a)
    side = P @ nrm.T - off                                   # > 0 outside a face plane
    inside = np.all(side < 0, axis=1)
    clear = np.min(np.abs(side) / np.linalg.norm(nrm, axis=1), axis=1) > 1e-9
    got = SC.classify(ext, cells, dist, counts) == SC.CONDUCTOR
    assert np.array_equal(got[clear], inside[clear]) and inside.sum() > 50
    for d in range(6):
        ax, sg = d // 2, (1.0 if d % 2 == 0 else -1.0)
        e = np.zeros(3)
        e[ax] = sg
        den = nrm @ e
        with np.errstate(divide="ignore", invalid="ignore"):
            tt = (off[None, :] - P @ nrm.T) / den[None, :]       # distance along the ray to each plane
        with np.errstate(invalid="ignore"):
            hitp = P[:, None, :] + tt[..., None] * e
        # barycentric containment of the hit point in each triangle
        v0, v1 = b - a, c - a
        w = hitp - a[None]
        d00, d01, d11 = (v0 * v0).sum(1), (v0 * v1).sum(1), (v1 * v1).sum(1)
        d20, d21 = (w * v0[None]).sum(2), (w * v1[None]).sum(2)
        den2 = d00 * d11 - d01 * d01
        bv = (d11 * d20 - d01 * d21) / den2
        bw = (d00 * d21 - d01 * d20) / den2
        with np.errstate(invalid="ignore"):
            ok = (bv >= -1e-12) & (bw >= -1e-12) & (bv + bw <= 1 + 1e-12) & (tt >= 0) & np.isfinite(tt)
        want = np.where(ok, tt, np.inf).min(axis=1)
        np.testing.assert_allclose(dist[clear, d], want[clear], rtol=1e-9, atol=1e-15)


def test_shortley_weller_rows_at_a_plane():
    """Next to the plane electrode x = c the x arm towards it is d = x - c (Shortley-Weller):
    diag 2/(h d) + 2/hy^2 + 2/hz^2, the +x neighbour -2/(h (h + d)), no -x entry; elsewhere the box stencil;
    without electrodes the operator is the box operator."""
    c = 0.0123
    v, t = box_mesh([-1.0, -1.0, -1.0], [c, 1.0, 1.0])
    dist, counts = SC.axis_ray_queries(v, t, EXT, CELLS)
    ct = SC.classify(EXT, CELLS, dist, counts)
    diag, off = SC.sw_coefficients(EXT, CELLS, ct, dist)
    h = [L / n for L, n in zip(EXT, CELLS)]
    x, y, z = _nodes(EXT, CELLS)
    first = x == np.min(x[x > c])
    d = x[first] - c
    np.testing.assert_allclose(diag[first], 2 / (h[0] * d) + 2 / h[1] ** 2 + 2 / h[2] ** 2, rtol=1e-14)
    np.testing.assert_allclose(off[first, 0], -2 / (h[0] * (h[0] + d)), rtol=1e-14)
    assert np.all(off[first, 1] == 0)
    far = x > np.min(x[x > c]) + h[0] / 2
    inner = far & (x < x.max())
    np.testing.assert_allclose(off[inner, 0], -1 / h[0] ** 2, rtol=1e-14)
    empty = SC.sw_operator(EXT, CELLS, np.zeros(x.size, dtype=np.int32), np.full((x.size, 6), np.inf))[0]
    assert abs(empty - SC.operator(EXT, CELLS)).max() == 0.0


def test_manufactured_solution_converges_second_order():
    """phi = sin(pi (x - c) / (X - c)) sin(pi (y + Y) / 2Y) sin(pi (z + Z) / 2Z) vanishes on the electrode
    face x = c and one spacing outside the box (X = Lx/2 + hx/2, ...); with rho = -eps0 lap(phi) the
    Shortley-Weller solution converges to it at second order."""
    c = 0.0123
    v, t = box_mesh([-1.0, -1.0, -1.0], [c, 1.0, 1.0])
    errs = []
    for k in (1, 2):
        cells = tuple(n * 2 * k for n in CELLS)
        h = [L / n for L, n in zip(EXT, cells)]
        X, Y, Z = [L / 2 + hh / 2 for L, hh in zip(EXT, h)]
        x, y, z = _nodes(EXT, cells)
        ex = np.sin(np.pi * (x - c) / (X - c)) * np.sin(np.pi * (y + Y) / (2 * Y)) * np.sin(np.pi * (z + Z) / (2 * Z))
        lam = (np.pi / (X - c)) ** 2 + (np.pi / (2 * Y)) ** 2 + (np.pi / (2 * Z)) ** 2
        dist, counts = SC.axis_ray_queries(v, t, EXT, cells)
        ct = SC.classify(EXT, cells, dist, counts)
        rho = SC.EPS0 * lam * ex
        phi = SC.solve_with_conductors(EXT, cells, rho.reshape(cells), ct, dist).ravel()
        act = ct != SC.CONDUCTOR
        assert np.all(phi[~act] == 0)
        errs.append(np.max(np.abs(phi[act] - ex[act])) / np.max(np.abs(ex[act])))
    print("manufactured-solution errors", errs)
    assert errs[0] < 0.05 and errs[0] / errs[1] > 3.0, errs
