"""pypatools_amd: MI355X-native particle push + field gather engine.

Drop-in for PyPATools' hot path (src/PyPATools/pusher.py, field.py,
field_src/interpolators.py, species.py): ``Pusher``, ``Field`` and
``IonSpecies`` with the reference's names and behaviour, backed by
hand-written gfx950 HIP kernels in libpypat_hip.so (C ABI: include/pypat_hip.h).
"""
from .constants import CLIGHT, ECHARGE, AMU_MEV, EMASS_MEV, PMASS_MEV, MUMASS_MEV  # noqa: F401
from .species import IonSpecies, PRESETS  # noqa: F401
from .field import (FieldBase, Field, CompositeField, ScaledField, FieldMap, RFField,  # noqa: F401
                    cartesian_to_cylindrical, cylindrical_to_cartesian, transform_cylindrical_field_to_cartesian)
from .interpolators import get_interpolator, HipInterpolator, DeviceFieldMap  # noqa: F401
from .pusher import Pusher  # noqa: F401
from .engine import Tracker, ParticleState  # noqa: F401

__version__ = "0.1.0"
