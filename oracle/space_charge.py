"""Space-charge field on a box mesh without electrodes (test infrastructure only).

Restates src/PyPATools/poisson_amg.py for a domain with no conductor cells:
mesh of cell-centre nodes (:227-239), CIC deposit of the CPU path (:849-911),
the 7-point operator with phi = 0 one spacing outside the box (:293-444), the
right-hand side rho / eps0 (:913-921) and E = -grad(phi) (:98-130).  The system
is solved with SciPy's sparse direct solver (the reference iterates GMRES + AMG
to a 1e-6 residual).
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
