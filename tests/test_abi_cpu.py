"""The C ABI library without a GPU: it loads, exports every symbol declared
in include/emfdtd.h, and its host-only closed forms agree with the oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "emfdtd.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^EMF_API\s+[\w\s\*]+?\b(emf_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "emf_solve" in names and "emf_update_H" in names and len(names) >= 40


def test_library_exports_every_declared_symbol():
    import ctypes
    from emopt_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared set
    assert sorted(_lib.exported_symbols()) == declared()


def test_host_closed_forms_match_oracle():
    from emopt_amd import _lib
    from oracle import driver as D
    L = _lib.load()._raw
    o = D.load(D.ORACLE_LIB)
    rng = np.random.default_rng(0)
    for _ in range(100):
        t = np.sort(rng.uniform(0, 50, 3)); f = rng.standard_normal(3)
        p1 = L.emf_calc_phase_3T(*t, *f); p2 = o.FDTD_calc_phase_3T(*t, *f)
        assert p1 == p2
        assert L.emf_calc_amplitude_3T(*t, *f, p1) == o.FDTD_calc_amplitude_3T(*t, *f, p2)
        assert L.emf_calc_phase_2T(t[0], t[1], f[0], f[1]) == o.FDTD_calc_phase_2T(t[0], t[1], f[0], f[1])
        p2t = L.emf_calc_phase_2T(t[0], t[1], f[0], f[1])
        assert (L.emf_calc_amplitude_2T(t[0], t[1], f[0], f[1], p2t)
                == o.FDTD_calc_amplitude_2T(t[0], t[1], f[0], f[1], p2t))


def test_error_path_without_handle():
    from emopt_amd import _lib
    L = _lib.load()
    with pytest.raises(_lib.EmfError):
        L.emf_set_dt(None, 1.0)
    assert "null" in L._raw.emf_last_error().decode()
