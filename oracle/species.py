"""IonSpecies restatement (q/m and masses).  Test infrastructure only.

Follows src/PyPATools/species.py:
  * preset table            species.py:7-36
  * rest mass = a * AMU_MEV species.py:117  (NOT the preset 'mass_mev' entry)
  * mass in kg              species.py:123  m_kg = m_MeV * e * 1e6 / c**2.0
  * q_over_m                species.py:158-163  q * e / m_kg
"""
from .constants import AMU_MEV, CLIGHT, ECHARGE, EMASS_MEV, MUMASS_MEV

# name -> (a, z, q); a in amu.  species.py:7-36
PRESETS = {
    "proton": (1.00727647, 1.0, 1.0),
    "H_1-": (1.00837361135, 1.0, 1.0),
    "electron": (EMASS_MEV / AMU_MEV, 0.0, -1.0),
    "muon": (MUMASS_MEV / AMU_MEV, 0.0, -1.0),
    "H2_1+": (2.01533, 2.0, 1.0),
    "4He_2+": (4.001506179127, 2.0, 2.0),
}


def mass_mev(name):
    a, _z, _q = PRESETS[name]
    return a * AMU_MEV  # species.py:117


def mass_kg(name):
    return mass_mev(name) * ECHARGE * 1.0e6 / CLIGHT ** 2.0  # species.py:123


def q_over_m(name):
    _a, _z, q = PRESETS[name]
    return q * ECHARGE / mass_kg(name)  # species.py:163
