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
  These are approximations of Chantler's values away from absorption edges.

f1/f2 only perturb f = Z + f1 + i f2 at the 1e-3 level for light elements at
10-13 keV; the values are an input of the path and their parity with xraydb
is unpinned.
"""

from __future__ import annotations

import json
import math
import os

# element: ((f1, f2) at Cu K-alpha, (f1, f2) at Mo K-alpha)
_E_CU = 8047.8
_E_MO = 17479.3
_BUILTIN = {
    'H': ((0.0, 0.0), (0.0, 0.0)),
    'C': ((0.0181, 0.0091), (0.0033, 0.0016)),
    'N': ((0.0311, 0.0180), (0.0061, 0.0033)),
    'O': ((0.0492, 0.0322), (0.0106, 0.0060)),
    'F': ((0.0727, 0.0534), (0.0171, 0.0103)),
    'Si': ((0.2541, 0.3302), (0.0817, 0.0704)),
    'S': ((0.3331, 0.5567), (0.1246, 0.1234)),
}


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
    """(f1, f2) of ``element`` at ``energy`` eV; KeyError if unknown."""
    path = os.environ.get("GIWAXSIM_F1F2_TABLE")
    if path:
        return _from_json(path, element, float(energy))
    (c1, c2), (m1, m2) = _BUILTIN[element]
    e = float(energy)
    return _powerlaw(e, c1, m1), _powerlaw(e, c2, m2)
