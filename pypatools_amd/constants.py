"""Physical constants (CODATA 2022, the values scipy.constants 1.15 provides).

Same names as src/PyPATools/global_variables.py:36-43 of the reference, which
reads them from scipy.constants; written out so the engine does not depend on
scipy at run time (tests check them against scipy.constants).
"""
CLIGHT = 299792458.0                 # m/s
ECHARGE = 1.602176634e-19            # C
EPS0 = 8.8541878188e-12              # F/m
AMU_MEV = 931.49410372               # MeV
AMU_KG = 1.66053906892e-27           # kg
EMASS_MEV = 0.51099895069            # MeV
PMASS_MEV = 938.27208943             # MeV
MUMASS_MEV = 105.6583755             # MeV

# unit factors, global_variables.py:46-58
kV, MV, GV = 1e3, 1e6, 1e9
kHz, MHz = 1e3, 1e6
km, dm, cm, mm = 1e3, 1e-1, 1e-2, 1e-3
pA, nA, uA, mA = 1e-12, 1e-9, 1e-6, 1e-3
