"""Space-charge field of a bunch on a box mesh (PyAMGPoissonSolver.solve,
src/PyPATools/poisson_amg.py:616-797), for domains without electrodes.

The charge density is deposited on the device straight from the particle
state (cloud in cell), the Poisson equation -lap(phi) = rho / eps0 for the
reference's 7-point operator with phi = 0 one spacing outside the box is
solved directly by discrete sine transforms, and E = -grad(phi) is taken by
central differences.  The result is a Field on the cell-centre grid, ready to
be added to the external fields of a push.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray


class SpaceChargeSolver:
    """Box mesh of mesh_cells = (nx, ny, nz) cell-centre nodes spanning domain_extent = (Lx, Ly, Lz) [m]."""

    def __init__(self, domain_extent, mesh_cells):
        self.extent = tuple(float(v) for v in domain_extent)
        self.cells = tuple(int(v) for v in mesh_cells)
        h = ctypes.c_void_p()
        _lib.call("pp_poisson_create", *self.cells, *self.extent, ctypes.byref(h))
        self.handle = h.value
        self.h = tuple(L / n for L, n in zip(self.extent, self.cells))
        n = int(np.prod(self.cells))
        self.rho = DeviceArray((n,))
        self.phi = DeviceArray((n,))
        self.E = [DeviceArray((n,)) for _ in range(3)]
        self.turn_count = 0

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
        _lib.call("pp_poisson_solve", self.handle, self.rho.ptr, self.phi.ptr, None)
        _lib.call("pp_efield_from_potential", self.handle, self.phi.ptr, *(e.ptr for e in self.E), None)
        self.turn_count += 1

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
