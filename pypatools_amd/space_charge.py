"""Space-charge field of a bunch on a box mesh (PyAMGPoissonSolver,
src/PyPATools/poisson_amg.py:133-797), with or without electrodes.

The charge density is deposited on the device straight from the particle
state (cloud in cell), and E = -grad(phi) is taken by central differences.  The
result is a Field on the cell-centre grid, ready to be added to the external
fields of a push.

- Without electrodes, -lap(phi) = rho / eps0 for the reference's 7-point
  operator with phi = 0 one spacing outside the box is solved directly by
  discrete sine transforms.
- With an electrode assembly (closed surfaces are solid, grounded conductors), the
  nodes are classified from crossings along their six axis rays (poisson_amg.py:241-291).
  The non-conductor nodes carry the reference's Shortley-Weller rows (:293-444),
  and the system is solved on the device by BiCGSTAB preconditioned with the box
  solve.  A TriangleElectrodes mesh is queried by the library itself.  An object
  with py_electrodes' compute_axis_aligned_surface_intersections(nodes, axes='all',
  use_gpu=...) -> (distances (N, 6), counts (N, 6)) is asked for its ray data, as
  the reference asks it.
"""
import ctypes
import warnings

import numpy as np

from . import _lib
from .device import DeviceArray


INTERIOR, BOUNDARY, CONDUCTOR = 0, 1, 2       # poisson_amg.py:86-90


class SpaceChargeSolver:
    """Box mesh of mesh_cells = (nx, ny, nz) cell-centre nodes spanning domain_extent = (Lx, Ly, Lz) [m],
    optionally bounded by an electrode assembly (grounded conductors)."""

    def __init__(self, domain_extent, mesh_cells, electrode_assembly=None, solver_tol=1e-10, max_iterations=500):
        self.extent = tuple(float(v) for v in domain_extent)
        self.cells = tuple(int(v) for v in mesh_cells)
        h = ctypes.c_void_p()
        _lib.call("pp_poisson_create", *self.cells, *self.extent, ctypes.byref(h))
        self.handle = h.value
        self.h = tuple(L / n for L, n in zip(self.extent, self.cells))
        n = int(np.prod(self.cells))
        self.n_dofs = n
        self.rho = DeviceArray((n,))
        self.phi = DeviceArray((n,))
        self.E = [DeviceArray((n,)) for _ in range(3)]
        self.turn_count = 0
        self.electrodes = electrode_assembly
        self.solver_tol = float(solver_tol)
        _lib.call("pp_poisson_set_tolerance", self.handle, float(solver_tol), int(max_iterations))
        if electrode_assembly is not None:
            self._set_electrodes(electrode_assembly)

    def _set_electrodes(self, assy):
        if hasattr(assy, "vertices") and hasattr(assy, "triangles"):
            v = np.ascontiguousarray(assy.vertices, dtype=np.float64).reshape(-1, 3)
            t = np.ascontiguousarray(assy.triangles, dtype=np.int32).reshape(-1, 3)
            _lib.call("pp_poisson_set_conductors", self.handle, v.ctypes.data, len(v), t.ctypes.data, len(t))
        elif hasattr(assy, "compute_axis_aligned_surface_intersections"):
            # poisson_amg.py:250-251: the assembly answers the ray queries itself
            dist, counts = assy.compute_axis_aligned_surface_intersections(self.mesh_nodes(), axes="all",
                                                                          use_gpu=False)
            dist = np.ascontiguousarray(dist, dtype=np.float64).reshape(self.n_dofs, 6)
            counts = np.ascontiguousarray(counts, dtype=np.int32).reshape(self.n_dofs, 6)
            _lib.call("pp_poisson_set_ray_data", self.handle, dist.ctypes.data, counts.ctypes.data)
        else:
            raise TypeError("electrode_assembly needs vertices/triangles (a triangle mesh) or "
                            "compute_axis_aligned_surface_intersections (py_electrodes)")

    def mesh_nodes(self):
        """poisson_amg.py:227-239: (N, 3) cell-centre node coordinates, z fastest."""
        g = self.grid()
        xx, yy, zz = np.meshgrid(g["x"], g["y"], g["z"], indexing="ij")
        return np.column_stack([xx.ravel(), yy.ravel(), zz.ravel()])

    def conductor_data(self):
        """(cell_type (N,), boundary_distances (N, 6), hit_counts (N, 6)) of the electrode classification."""
        if self.electrodes is None:
            raise ValueError("no electrode assembly")
        t = np.zeros(self.n_dofs, dtype=np.int32)
        d = np.zeros((self.n_dofs, 6))
        c = np.zeros((self.n_dofs, 6), dtype=np.int32)
        _lib.call("pp_poisson_conductor_data", self.handle, t.ctypes.data, d.ctypes.data, c.ctypes.data)
        return t, d, c

    @property
    def cell_type(self):
        return self.conductor_data()[0]

    @property
    def boundary_distances(self):
        return self.conductor_data()[1]

    def solve_stats(self):
        """(iterations, relative residual) of the last solve (0 iterations: the direct box solve)."""
        it, res = ctypes.c_int32(), ctypes.c_double()
        _lib.call("pp_poisson_solve_stats", self.handle, ctypes.byref(it), ctypes.byref(res))
        return int(it.value), float(res.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                _lib.lib().pp_poisson_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self.handle = None

    def grid(self):
        """poisson_amg.py:743-752: cell-centre axes."""
        return {a: np.linspace(-L / 2 + h / 2, L / 2 - h / 2, n)
                for a, L, h, n in zip("xyz", self.extent, self.h, self.cells)}

    def _solve_device(self, x, y, z, q_ptr, q_const, n, alive_ptr=None):
        self.rho.fill(0.0)
        _lib.call("pp_cic_deposit", self.handle, x, y, z, q_ptr, float(q_const), alive_ptr, int(n), self.rho.ptr, None)
        self._solve_rho()

    def _solve_rho(self):
        _lib.call("pp_poisson_solve", self.handle, self.rho.ptr, self.phi.ptr, None)
        _lib.call("pp_efield_from_potential", self.handle, self.phi.ptr, *(e.ptr for e in self.E), None)
        self.turn_count += 1
        if self.electrodes is not None:
            it, res = self.solve_stats()
            if not res <= self.solver_tol:      # poisson_amg.py:711-715 warns when GMRES stops short
                warnings.warn(f"BiCGSTAB did not converge: relative residual {res:.2e} after {it} iterations "
                              f"(tolerance {self.solver_tol:.1e})", RuntimeWarning)

    def solve_density(self, rho):
        """phi (nx, ny, nz) for a given charge density rho [C/m^3] on the nodes (and the field in self.E)."""
        self.rho.upload(np.ascontiguousarray(rho, dtype=np.float64).reshape(-1))
        self._solve_rho()
        return self.phi.to_host().reshape(self.cells)

    def _field(self):
        from .field import Field
        vals = {a: e.to_host().reshape(self.cells) for a, e in zip("xyz", self.E)}
        return Field.from_arrays(self.grid(), vals, label=f"Space-charge E-field (turn {self.turn_count})", dim=3)

    def solve(self, particles, charges):
        """poisson_amg.py:616-797: (phi_3d, E_field) for positions (N, 3) [m] and charges (N,) [C]."""
        p = np.ascontiguousarray(np.asarray(particles, dtype=np.float64).reshape(-1, 3).T)
        q = np.ascontiguousarray(charges, dtype=np.float64).reshape(-1)
        n = p.shape[1]
        if q.shape[0] != n:
            raise ValueError("particles and charges differ in length")
        dp = DeviceArray.from_host(p) if n else DeviceArray((3,))
        dq = DeviceArray.from_host(q) if n else DeviceArray((1,))
        self._solve_device(dp.view_ptr(0), dp.view_ptr(n), dp.view_ptr(2 * n), dq.ptr, 0.0, n)
        return self.phi.to_host().reshape(self.cells), self._field()

    def solve_state(self, state, charge_per_particle, alive=None):
        """The same from a device-resident ParticleState with one charge per particle (no host copy of the state)."""
        self._solve_device(state.x.ptr, state.y.ptr, state.z.ptr, None, charge_per_particle, state.n,
                           alive.ptr if alive is not None else None)
        return self.phi.to_host().reshape(self.cells), self._field()
