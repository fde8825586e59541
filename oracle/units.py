"""Unit constants, CODATA 2014, derived the way ase.units derives them.

Reference: ``amaterasu/units.py:55-80`` takes ``Bohr`` and ``Ha`` from ``ase.units``
(ase ships CODATA 2014). ase is absent here, so the same derivation from the
CODATA 2014 fundamental constants is restated.
"""

from __future__ import annotations

import math

_C = 299792458.0
_MU0 = 4.0e-7 * math.pi
_HPLANCK = 6.626070040e-34
_E = 1.6021766208e-19
_ME = 9.10938356e-31
_EPS0 = 1.0 / _MU0 / _C**2
_HBAR = _HPLANCK / (2.0 * math.pi)

#: Angstrom per Bohr (ase: 4e10 pi eps0 hbar^2 / (me e^2)).
BOHR_IN_ANG = 4.0e10 * math.pi * _EPS0 * _HBAR**2 / _ME / _E**2
ANG_IN_BOHR = 1.0 / BOHR_IN_ANG
#: eV per Hartree (ase: me e^3 / (16 pi^2 eps0^2 hbar^2)).
HARTREE_IN_EV = _ME * _E**3 / 16.0 / math.pi**2 / _EPS0**2 / _HBAR**2
EV_IN_HARTREE = 1.0 / HARTREE_IN_EV
RYDBERG_IN_HARTREE = 0.5
