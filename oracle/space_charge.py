"""Space-charge field on a box mesh, with or without electrodes (test infrastructure only).

Restates src/PyPATools/poisson_amg.py: mesh of cell-centre nodes (:227-239), CIC
deposit of the CPU path (:849-911), the 7-point operator with phi = 0 one spacing
outside the box (:293-444), the right-hand side rho / eps0 (:913-921) and
E = -grad(phi) (:98-130).  With electrodes: node classification (:241-291) and the
Shortley-Weller operator over the non-conductor nodes (:293-444).  The systems are
solved with SciPy's sparse direct solver (the reference iterates GMRES + AMG to a
1e-6 residual).

The reference gets crossing counts and distances along the six axis rays of every
node from py_electrodes' compute_axis_aligned_surface_intersections, a package
absent here; `axis_ray_queries` states the rule this build uses in its place (its
agreement with py_electrodes is parity unpinned; it is pinned by geometry).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS0 = 8.8541878188e-12      # global_variables.py:38 (scipy.constants.epsilon_0)


def mesh_axes(extent, cells):
    """poisson_amg.py:227-239: cell-centre coordinates of each axis."""
    out = []
    for L, n in zip(extent, cells):
        h = L / n
        out.append(np.linspace(-L / 2 + h / 2, L / 2 - h / 2, n))
    return out


def cic_deposit(extent, cells, pos, charges):
    """poisson_amg.py:849-911 (CPU path): rho [C/m^3] on the (nx, ny, nz) nodes."""
    nx, ny, nz = cells
    h = [L / n for L, n in zip(extent, cells)]
    rho = np.zeros(nx * ny * nz)
    g = []
    for a in range(3):
        u = (pos[:, a] + extent[a] / 2) / h[a]
        g.append(np.clip(u, 0, cells[a] - 1))
    ix = [np.floor(u).astype(np.int64) for u in g]
    f = [g[a] - ix[a] for a in range(3)]
    ix = [np.clip(ix[a], 0, cells[a] - 2) for a in range(3)]
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                jx, jy, jz = (ix[0] + dx) % nx, (ix[1] + dy) % ny, (ix[2] + dz) % nz
                idx = jx * (ny * nz) + jy * nz + jz
                w = ((1.0 - f[0] if dx == 0 else f[0]) * (1.0 - f[1] if dy == 0 else f[1])) * \
                    (1.0 - f[2] if dz == 0 else f[2])
                np.add.at(rho, idx, charges * w)
    return (rho / (h[0] * h[1] * h[2])).reshape(nx, ny, nz)


def operator(extent, cells):
    """poisson_amg.py:293-444 with no conductor cells: diag 2/h^2 per axis, neighbours -1/h^2."""
    nx, ny, nz = cells
    h = [L / n for L, n in zip(extent, cells)]
    N = nx * ny * nz
    idx = np.arange(N).reshape(nx, ny, nz)
    rows, cols, vals = [np.arange(N)], [np.arange(N)], [np.full(N, sum(2.0 / (hh * hh) for hh in h))]
    for a, hh in enumerate(h):
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[a] = slice(0, cells[a] - 1)
        sl_hi[a] = slice(1, cells[a])
        lo, hi = idx[tuple(sl_lo)].ravel(), idx[tuple(sl_hi)].ravel()
        off = -2.0 / (hh * (hh + hh))
        rows += [lo, hi]
        cols += [hi, lo]
        vals += [np.full(lo.size, off), np.full(lo.size, off)]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def solve(extent, cells, rho):
    """phi with A phi = rho / eps0."""
    A = operator(extent, cells)
    return spla.spsolve(A.tocsc(), rho.ravel() / EPS0).reshape(cells)


def gradient(extent, cells, phi):
    """poisson_amg.py:98-130: E = -grad(phi) on interior nodes, 0 on the outer layer."""
    h = [L / n for L, n in zip(extent, cells)]
    E = [np.zeros(phi.shape) for _ in range(3)]
    c = (slice(1, -1),) * 3
    E[0][c] = -(phi[2:, 1:-1, 1:-1] - phi[:-2, 1:-1, 1:-1]) / (2 * h[0])
    E[1][c] = -(phi[1:-1, 2:, 1:-1] - phi[1:-1, :-2, 1:-1]) / (2 * h[1])
    E[2][c] = -(phi[1:-1, 1:-1, 2:] - phi[1:-1, 1:-1, :-2]) / (2 * h[2])
    return E


# --------------------------------------------------------------------------- electrode-bounded domains

INTERIOR, BOUNDARY, CONDUCTOR = 0, 1, 2       # poisson_amg.py:86-90


def _owned(du, dv):
    """Top-left rule: a counter-clockwise triangle owns the points on its edge of direction (du, dv) when
    dv > 0, or dv == 0 and du < 0.  Opposite directions never both qualify, so a point on an edge shared by
    two triangles of one surface belongs to exactly one of them."""
    return (dv > 0) | ((dv == 0) & (du < 0))


def axis_ray_queries(vertices, triangles, extent, cells):
    """Crossings of the six axis rays of every node with a triangle mesh: (dist (N, 6), counts (N, 6)) for the
    directions (+x, -x, +y, -y, +z, -z), nodes in (nx, ny, nz) order; dist is inf where a ray meets nothing.

    For axis a the mesh is projected along a onto the plane of the other two axes (u, v) = ((a+1)%3, (a+2)%3),
    relative to the node's line: P_i = (V_i[u] - o_u, V_i[v] - o_v).  Edge functions E_ij = P_i.u P_j.v -
    P_i.v P_j.u are exact negatives for the two triangles sharing an edge.  A triangle with a non-zero
    projected area A = (E_01 + E_12) + E_20 crosses the line when, oriented counter-clockwise (sign s of A),
    every s E_ij is > 0, or == 0 with the edge (direction s (P_j - P_i)) owned (`_owned`).  The crossing lies
    at depth (E_12 V_0[a] + E_20 V_1[a] + E_01 V_2[a]) / A; a node at coordinate x counts it on its + ray
    when depth >= x (distance depth - x) and on its - ray otherwise (distance x - depth).
    """
    V = np.asarray(vertices, dtype=np.float64)
    T = np.asarray(triangles, dtype=np.int64).reshape(-1, 3)
    nx, ny, nz = cells
    axes = mesh_axes(extent, cells)
    dist = np.full((nx, ny, nz, 6), np.inf)
    counts = np.zeros((nx, ny, nz, 6), dtype=np.int64)
    if len(T) == 0:
        return dist.reshape(-1, 6), counts.reshape(-1, 6)
    for a in range(3):
        u, v = (a + 1) % 3, (a + 2) % 3
        ou, ov = np.meshgrid(axes[u], axes[v], indexing="ij")            # lines (n_u, n_v)
        ou, ov = ou.ravel()[:, None], ov.ravel()[:, None]
        Pu = [V[T[:, i], u][None, :] - ou for i in range(3)]            # (lines, triangles)
        Pv = [V[T[:, i], v][None, :] - ov for i in range(3)]
        E01 = Pu[0] * Pv[1] - Pv[0] * Pu[1]
        E12 = Pu[1] * Pv[2] - Pv[1] * Pu[2]
        E20 = Pu[2] * Pv[0] - Pv[2] * Pu[0]
        A = (E01 + E12) + E20
        sg = np.sign(A)
        inside = A != 0
        for E, i, j in ((E01, 0, 1), (E12, 1, 2), (E20, 2, 0)):
            du = sg * (V[T[:, j], u] - V[T[:, i], u])[None, :]
            dv = sg * (V[T[:, j], v] - V[T[:, i], v])[None, :]
            se = sg * E
            inside &= (se > 0) | ((se == 0) & _owned(du, dv))
        with np.errstate(invalid="ignore", divide="ignore"):
            depth = ((E12 * V[T[:, 0], a][None, :] + E20 * V[T[:, 1], a][None, :]) + E01 * V[T[:, 2], a][None, :]) / A
        xs = axes[a]
        for line in range(ou.shape[0]):
            d = np.sort(depth[line][inside[line]])
            if d.size == 0:
                continue
            iu, iv = divmod(line, len(axes[v]))
            for ia, x in enumerate(xs):
                ahead, behind = d[d >= x], d[d < x]
                idx = [0, 0, 0]
                idx[a], idx[u], idx[v] = ia, iu, iv
                cell = tuple(idx)
                counts[cell + (2 * a,)] = ahead.size
                counts[cell + (2 * a + 1,)] = behind.size
                if ahead.size:
                    dist[cell + (2 * a,)] = np.min(ahead - x)
                if behind.size:
                    dist[cell + (2 * a + 1,)] = np.min(x - behind)
    return dist.reshape(-1, 6), counts.reshape(-1, 6)


def classify(extent, cells, dist, counts):
    """poisson_amg.py:241-291: CONDUCTOR where a ray has an odd crossing count, BOUNDARY where a distance is
    at most the spacing of its axis (compared with the spacings rounded to float32, as there), else INTERIOR."""
    h = [L / n for L, n in zip(extent, cells)]
    is_conductor = (counts % 2 != 0).any(axis=1)
    sizes = np.array([h[0], h[0], h[1], h[1], h[2], h[2]], dtype=np.float32)
    is_boundary = (dist <= sizes).any(axis=1) & ~is_conductor
    t = np.full(dist.shape[0], INTERIOR, dtype=np.int32)
    t[is_boundary] = BOUNDARY
    t[is_conductor] = CONDUCTOR
    return t


def sw_coefficients(extent, cells, cell_type, dist):
    """poisson_amg.py:293-444: Shortley-Weller row of every node as (diag (N,), off (N, 6)) for the neighbours
    (+x, -x, +y, -y, +z, -z).  Towards a conductor neighbour the arm is max(distance, 1e-3 h) and carries no
    entry; towards the outside of the box it is h with no entry; conductor rows are left at 0."""
    nx, ny, nz = cells
    h = [L / n for L, n in zip(extent, cells)]
    N = nx * ny * nz
    t = np.asarray(cell_type).reshape(nx, ny, nz)
    ijk = np.indices(cells)
    diag = np.zeros((nx, ny, nz))
    off = np.zeros((nx, ny, nz, 6))
    D = np.asarray(dist).reshape(nx, ny, nz, 6)
    for a in range(3):
        n = cells[a]
        has_p = ijk[a] < n - 1
        has_m = ijk[a] > 0
        tp = np.full(cells, INTERIOR)
        tm = np.full(cells, INTERIOR)
        sl_dst, sl_src = [slice(None)] * 3, [slice(None)] * 3
        sl_dst[a], sl_src[a] = slice(0, n - 1), slice(1, n)
        tp[tuple(sl_dst)] = t[tuple(sl_src)]
        tm[tuple(sl_src)] = t[tuple(sl_dst)]
        cond_p = has_p & (tp == CONDUCTOR)
        cond_m = has_m & (tm == CONDUCTOR)
        hp = np.where(cond_p, np.maximum(D[..., 2 * a], h[a] * 1e-3), h[a])
        hm = np.where(cond_m, np.maximum(D[..., 2 * a + 1], h[a] * 1e-3), h[a])
        diag = diag + 2.0 / (hp * hm)
        off[..., 2 * a] = np.where(has_p & ~cond_p, -2.0 / (hp * (hp + hm)), 0.0)
        off[..., 2 * a + 1] = np.where(has_m & ~cond_m, -2.0 / (hm * (hp + hm)), 0.0)
    active = t != CONDUCTOR
    diag = np.where(active, diag, 0.0)
    off = np.where(active[..., None], off, 0.0)
    return diag.reshape(N), off.reshape(N, 6)


def sw_operator(extent, cells, cell_type, dist):
    """The reduced Shortley-Weller matrix over the non-conductor nodes and their indices."""
    nx, ny, nz = cells
    N = nx * ny * nz
    diag, off = sw_coefficients(extent, cells, cell_type, dist)
    keep = np.where(np.asarray(cell_type) != CONDUCTOR)[0]
    new = np.full(N, -1, dtype=np.int64)
    new[keep] = np.arange(keep.size)
    strides = [(ny * nz), (ny * nz), nz, nz, 1, 1]
    signs = [1, -1, 1, -1, 1, -1]
    rows, cols, vals = [new[keep]], [new[keep]], [diag[keep]]
    for d in range(6):
        o = off[keep, d]
        m = o != 0.0
        nb = keep[m] + signs[d] * strides[d]
        rows.append(new[keep[m]])
        cols.append(new[nb])
        vals.append(o[m])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(keep.size, keep.size))
    return A, keep


def solve_with_conductors(extent, cells, rho, cell_type, dist):
    """poisson_amg.py:616-735: phi on the full grid, 0 on conductor nodes."""
    A, keep = sw_operator(extent, cells, cell_type, dist)
    phi = np.zeros(int(np.prod(cells)))
    if keep.size:
        phi[keep] = spla.spsolve(A.tocsc(), rho.ravel()[keep] / EPS0)
    return phi.reshape(cells)
