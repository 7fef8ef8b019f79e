"""CPU oracle for the particle-push + trilinear-gather path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``pypatools_amd``)
imports, links or executes anything under ``oracle/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or as the timed CPU baseline), never as the measured or
shipped path.

The oracle is a NumPy FP64 restatement of PyPATools' pushers and field gather:

* ``constants``  -- physical constants (src/PyPATools/global_variables.py:36-43)
* ``species``    -- IonSpecies mass / q-over-m (src/PyPATools/species.py:7-187)
* ``kernels``    -- per-step integrator kernels, two execution forms:
    - "as-shipped" scalar form: one particle at a time with NumPy scalar
      arithmetic, exactly as the reference's no-numba path runs them
      (src/PyPATools/pusher.py:755-775); ``x**2`` is libm ``pow``.
    - canonical vectorised form over particles; ``x**2`` is ``x*x`` (what the
      Numba build of the reference compiles it to and what ``np.sum(v**2)``
      does).  This form is the definition the strict FP64 GPU kernels match
      bit for bit.
* ``interp``     -- trilinear/bilinear/linear gather with NumbaInterpolator
  semantics (src/PyPATools/field_src/interpolators.py:44-135) and the
  CoordinateMapper path (interpolators.py:415-526).
* ``field``      -- Field dispatch, scaling, ScaledField/CompositeField
  (src/PyPATools/field.py:213-244, 441-504, 595-690) plus the extensions the
  engine adds (periodic-z wrap, RF time dependence).
* ``pusher``     -- Pusher.push / push_batch / track / track_batch orchestration
  (src/PyPATools/pusher.py:430-1037).

Pinning status of each function is listed in ``oracle/PINNING.md``.
"""
