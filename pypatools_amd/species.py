"""IonSpecies: mass and charge of the tracked species.

Mirrors src/PyPATools/species.py (PRESETS :7-36, IonSpecies :39-187): the rest
mass is ``a * AMU_MEV`` (species.py:117), the mass in kg
``m_MeV * e * 1e6 / c**2.0`` (:123) and ``q_over_m = q * e / m_kg`` (:163).
"""
from .constants import AMU_MEV, CLIGHT, ECHARGE, EMASS_MEV, MUMASS_MEV, PMASS_MEV

PRESETS = {
    "proton": {"latex_label": r"$\mathrm{p}^+$", "mass_mev": PMASS_MEV, "a": 1.00727647, "z": 1.0, "q": 1.0},
    "H_1-": {"latex_label": r"$\mathrm{H}^1-$", "mass_mev": 1.00837361135 * AMU_MEV, "a": 1.00837361135,
             "z": 1.0, "q": 1.0},
    "electron": {"latex_label": r"$\mathrm{e}^-$", "mass_mev": EMASS_MEV, "a": EMASS_MEV / AMU_MEV, "z": 0.0,
                 "q": -1.0},
    "muon": {"latex_label": r"$\mu^-$", "mass_mev": MUMASS_MEV, "a": MUMASS_MEV / AMU_MEV, "z": 0.0, "q": -1.0},
    "H2_1+": {"latex_label": r"$\mathrm{H}_2^+$", "mass_mev": 2.01533 * AMU_MEV, "a": 2.01533, "z": 2.0, "q": 1.0},
    "4He_2+": {"latex_label": r"$^4\mathrm{He}^{2+}$", "mass_mev": 4.001506179127 * AMU_MEV,
               "a": 4.001506179127, "z": 2.0, "q": 2.0},
}


class IonSpecies:
    """Ion species by preset name, or by name plus a (amu), z and q."""

    def __init__(self, name, label=None, latex_label=None, a=None, z=None, q=None):
        if name in PRESETS:
            sp = PRESETS[name]
            label = name if label is None else label
            latex_label = sp["latex_label"] if latex_label is None else latex_label
            a, z, q = sp["a"], sp["z"], sp["q"]
        elif None in (a, z, q):
            # the reference prints and calls exit(1) here (species.py:91-97)
            raise ValueError(f"ion species {name} was initialized with missing values: a={a}, z={z}, q={q}")
        self._name = name
        self._label = name if label is None else label
        self._latex_label = name if latex_label is None else latex_label
        self._mass_mev = a * AMU_MEV
        self._a = a
        self._z = z
        self._q = q
        self._mass_kg = self._mass_mev * ECHARGE * 1.0e6 / CLIGHT ** 2.0

    def __str__(self):
        return (f"Ion Species {self._name} with label {self._label}:\n"
                f"A = {self._a}, Z = {self._z}, q = {self._q}\n"
                f"M_0 = {self._mass_mev} MeV/c^2 = {self._mass_kg} kg, Q = {self._q * ECHARGE} C\n"
                f"(For OPERA: M_0 = {self._mass_mev / EMASS_MEV} * M_electron)")

    @property
    def mass_electron_multiple(self):
        return self._mass_mev / EMASS_MEV

    @property
    def mass_mev(self):
        return self._mass_mev

    @property
    def mass_kg(self):
        return self._mass_kg

    @property
    def mass_opera(self):
        return self._mass_mev / EMASS_MEV

    @property
    def q_over_a(self):
        return self._q / self._a

    @property
    def q_over_m(self):
        return self._q * ECHARGE / self._mass_kg

    @property
    def label(self):
        return self._label

    @property
    def latex_label(self):
        return self._latex_label

    @property
    def name(self):
        return self._name

    @property
    def q(self):
        return self._q

    @property
    def charge(self):
        return self._q * ECHARGE

    @property
    def a(self):
        return self._a
