"""Offline anomalous scattering factors f' (f1) and f'' (f2).

The reference looks them up with xraydb's Chantler tables
(tools/utilities.py:342-366, ``xraydb.f1_chantler`` / ``f2_chantler``).
xraydb is not available here, so this module provides an offline provider:

* ``GIWAXSIM_F1F2_TABLE`` may name a JSON file ``{"C": [[E_eV, f1, f2], ...], ...}``
  (e.g. exported once from xraydb); values are interpolated linearly in
  log(E) between the listed energies.
* Otherwise a built-in table is used: Cromer-Liberman f', f'' at Cu K-alpha
  (8047.8 eV) and Mo K-alpha (17479.3 eV) from International Tables Vol. C,
  Table 4.2.6.8, interpolated as a power law in E between (and beyond) them.
  It holds the light elements H, B, C, N, O, F, Na, Mg, Al, Si, P, S, Cl, K and
  Ca, whose K edges lie below both lines (Ca's at 4.04 keV), so the power law
  does not cross an edge; these are approximations of Chantler's values.
* Any other element of the periodic table (H..U, the reference's ``ptable``)
  gets f1 = f2 = 0, i.e. f = Z, with a warning naming the element: heavier
  elements have edges between or near the two lines (Br's K edge is at
  13.47 keV), where no interpolation of two values is meaningful; give their
  values through ``GIWAXSIM_F1F2_TABLE``.  A symbol outside the periodic table
  raises KeyError.

f1/f2 only perturb f = Z + f1 + i f2 at the 1e-3 to 1e-2 level for the light
elements at 10-13 keV; the values are an input of the path and their parity
with xraydb is unpinned.
"""

from __future__ import annotations

import json
import math
import os
import warnings

from .ptable_dict import ptable

# element: ((f1, f2) at Cu K-alpha, (f1, f2) at Mo K-alpha)
_E_CU = 8047.8
_E_MO = 17479.3
_BUILTIN = {
    'H': ((0.0, 0.0), (0.0, 0.0)),
    'B': ((0.0090, 0.0039), (0.0013, 0.0007)),
    'C': ((0.0181, 0.0091), (0.0033, 0.0016)),
    'N': ((0.0311, 0.0180), (0.0061, 0.0033)),
    'O': ((0.0492, 0.0322), (0.0106, 0.0060)),
    'F': ((0.0727, 0.0534), (0.0171, 0.0103)),
    'Na': ((0.1353, 0.1239), (0.0362, 0.0249)),
    'Mg': ((0.1719, 0.1771), (0.0486, 0.0363)),
    'Al': ((0.2130, 0.2455), (0.0645, 0.0514)),
    'Si': ((0.2541, 0.3302), (0.0817, 0.0704)),
    'P': ((0.2955, 0.4335), (0.1023, 0.0942)),
    'S': ((0.3331, 0.5567), (0.1246, 0.1234)),
    'Cl': ((0.3639, 0.7018), (0.1484, 0.1585)),
    'K': ((0.3868, 1.0657), (0.2009, 0.2494)),
    'Ca': ((0.3641, 1.2855), (0.2262, 0.3064)),
}
_WARNED: set = set()


def _powerlaw(e, v_cu, v_mo):
    if v_cu == 0.0 or v_mo == 0.0:
        return v_cu + (v_mo - v_cu) * (e - _E_CU) / (_E_MO - _E_CU)
    p = math.log(v_mo / v_cu) / math.log(_E_MO / _E_CU)
    return v_cu * (e / _E_CU) ** p


def _from_json(path, element, energy):
    with open(path) as fh:
        tab = json.load(fh)
    rows = sorted(tab[element])
    es = [r[0] for r in rows]
    if energy <= es[0]:
        return rows[0][1], rows[0][2]
    if energy >= es[-1]:
        return rows[-1][1], rows[-1][2]
    for (e0, a0, b0), (e1, a1, b1) in zip(rows, rows[1:]):
        if e0 <= energy <= e1:
            t = (math.log(energy) - math.log(e0)) / (math.log(e1) - math.log(e0))
            return a0 + t * (a1 - a0), b0 + t * (b1 - b0)
    raise KeyError(element)


def f1_f2(element: str, energy: float):
    """(f1, f2) of ``element`` at ``energy`` eV: from the JSON table when one is named
    (KeyError for an element it lacks), else the built-in table, else (0, 0) with a
    warning for an element of the periodic table; KeyError for any other symbol."""
    path = os.environ.get("GIWAXSIM_F1F2_TABLE")
    if path:
        return _from_json(path, element, float(energy))
    if element not in _BUILTIN:
        if element not in ptable:
            raise KeyError(element)
        if element not in _WARNED:
            _WARNED.add(element)
            warnings.warn(f"no offline f1/f2 for '{element}': using f = Z (f1 = f2 = 0); "
                          "set GIWAXSIM_F1F2_TABLE to a JSON table of its values", stacklevel=2)
        return 0.0, 0.0
    (c1, c2), (m1, m2) = _BUILTIN[element]
    e = float(energy)
    return _powerlaw(e, c1, m1), _powerlaw(e, c2, m2)
