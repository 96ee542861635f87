"""Offline anomalous scattering factors (giwaxsim_amd.tools.f1f2), CPU only.

The reference reads f1/f2 from xraydb's Chantler tables (tools/utilities.py:342-366),
which are absent here, so parity of the values is unpinned.  What is pinned:

* the built-in light-element table against the physics that fixes its shape: f'' is
  proportional to E times the photoabsorption cross-section, which falls about as E^-3
  away from edges, so between Cu and Mo K-alpha f'' falls as E^p with p close to -2
  (a swapped line, a swapped f'/f'' column or a misplaced digit moves p far off), and
  f'' grows with Z along the table at either line;
* the provider's contract: element_table builds f = Z + f1 + i f2 for every element of
  the periodic table (the reference accepts any element xraydb knows), with a warning and
  f = Z where no offline value exists, and a KeyError for a symbol outside the table.
"""

import json
import math
import warnings

import numpy as np
import pytest

from giwaxsim_amd import engine
from giwaxsim_amd.tools import f1f2
from giwaxsim_amd.tools.ptable_dict import ptable


def test_builtin_table_follows_the_absorption_power_law():
    for el, ((c1, c2), (m1, m2)) in f1f2._BUILTIN.items():
        if el == 'H':
            continue
        p = math.log(m2 / c2) / math.log(f1f2._E_MO / f1f2._E_CU)
        assert -2.4 < p < -1.7, (el, p)
        # f' of these light elements falls with energy above their K edges
        assert m1 < c1, el
    order = sorted((ptable[e], e) for e in f1f2._BUILTIN if e != 'H')
    # the exponent drifts smoothly towards -1.8 as the K edge approaches the Cu line
    # (boron's four-decimal values are too coarse to take part)
    ps = [math.log(f1f2._BUILTIN[e][1][1] / f1f2._BUILTIN[e][0][1]) / math.log(f1f2._E_MO / f1f2._E_CU)
          for z, e in order if z >= 6]
    assert all(a < b for a, b in zip(ps, ps[1:])), ps
    for line in (0, 1):
        f2 = [f1f2._BUILTIN[e][line][1] for _, e in order]
        assert all(a < b for a, b in zip(f2, f2[1:])), (line, order)


def test_interpolation_hits_the_tabulated_lines():
    for el, (cu, mo) in f1f2._BUILTIN.items():
        assert f1f2.f1_f2(el, f1f2._E_CU) == pytest.approx(cu, rel=1e-12, abs=1e-15)
        assert f1f2.f1_f2(el, f1f2._E_MO) == pytest.approx(mo, rel=1e-12, abs=1e-15)


def test_element_table_light_elements_without_warning():
    els = np.array(['C', 'P', 'Cl', 'Na', 'C', 'Ca', 'P'])
    with warnings.catch_warnings():
        warnings.simplefilter('error')
        ids, ftab, counts, uniq = engine.element_table(els, 12700.0)
    assert list(uniq) == ['C', 'Ca', 'Cl', 'Na', 'P']
    assert list(counts) == [2, 1, 1, 1, 2]
    for i, e in enumerate(uniq):
        f1, f2 = f1f2.f1_f2(str(e), 12700.0)
        assert ftab[i] == complex(ptable[str(e)] + f1, f2)
        assert 0 < f1 < 1 and 0 < f2 < 1
    assert np.array_equal(uniq[ids], els)


def test_other_elements_fall_back_to_z_with_a_warning():
    f1f2._WARNED.discard('Br')
    with pytest.warns(UserWarning, match="no offline f1/f2 for 'Br'"):
        ids, ftab, _, uniq = engine.element_table(np.array(['Br', 'C']), 12700.0)
    assert ftab[list(uniq).index('Br')] == complex(35.0, 0.0)
    with warnings.catch_warnings():   # once per element
        warnings.simplefilter('error')
        assert f1f2.f1_f2('Br', 9000.0) == (0.0, 0.0)


def test_unknown_symbol_raises():
    with pytest.raises(KeyError):
        f1f2.f1_f2('Xx', 12700.0)
    with pytest.raises(KeyError):
        engine.element_table(np.array(['C', 'Xx']), 12700.0)


def test_json_table_overrides_and_is_strict(tmp_path, monkeypatch):
    path = tmp_path / "f1f2.json"
    path.write_text(json.dumps({"Br": [[10000.0, -1.0, 0.5], [14000.0, -3.0, 3.5]]}))
    monkeypatch.setenv("GIWAXSIM_F1F2_TABLE", str(path))
    f1, f2 = f1f2.f1_f2('Br', 12000.0)
    t = (math.log(12000.0) - math.log(10000.0)) / (math.log(14000.0) - math.log(10000.0))
    assert (f1, f2) == pytest.approx((-1.0 - 2.0 * t, 0.5 + 3.0 * t))
    with pytest.raises(KeyError):
        f1f2.f1_f2('C', 12000.0)
