"""Physical constants, taken from scipy.constants exactly as the reference does.

Follows src/PyPATools/global_variables.py:36-43 (CLIGHT, ECHARGE, AMU_MEV,
EMASS_MEV, PMASS_MEV, MUMASS_MEV from scipy.constants / CODATA).
Test infrastructure only (see oracle/__init__.py).
"""
from scipy import constants as _const

CLIGHT = _const.speed_of_light
ECHARGE = _const.elementary_charge
AMU_MEV = _const.value("atomic mass constant energy equivalent in MeV")
EMASS_MEV = _const.value("electron mass energy equivalent in MeV")
PMASS_MEV = _const.value("proton mass energy equivalent in MeV")
MUMASS_MEV = _const.value("muon mass energy equivalent in MeV")
