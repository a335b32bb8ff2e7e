"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed), and its host-side set-up matches the oracle."""
import ctypes
import os
import re

import numpy as np

from galic_amd import _abi, model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in ("galic_hip.h",):
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[a-zA-Z_][\w\s\*]*?\b([a-zA-Z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    lib = _abi.lib()
    decl = declared_functions()
    assert len(decl) >= 25
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_abi.EXPORTS) == decl


def test_version_and_error_channel():
    lib = _abi.lib()
    assert lib.galic_version() == 10100
    p = _abi.GalicParams()
    rc = lib.galic_grid_create(ctypes.byref(p), 1.5, 0.5)
    assert rc == -1
    assert b"fractions" in lib.galic_last_error()


def test_params_struct_layout_matches_c():
    # the C struct ends with two doubles after four int/double groups; its size is fixed by the header
    assert ctypes.sizeof(_abi.GalicParams) % 8 == 0
    p = model.make_params(model.config("halo-1e5"))
    assert p.DG_Fac > 1 and p.FG_Fac > 1 and p.Rmax > 0


def test_density_geometry_host_entry():
    p = model.make_params(model.config("halo-1e5"))
    nb = 1 << p.DG_MaxLevel
    cs = np.zeros(nb * nb)
    dist = np.zeros(model.nstack(p.DG_MaxLevel))
    _abi.check(_abi.lib().galic_density_geometry(ctypes.byref(p), _abi.dptr(cs), None, _abi.dptr(dist)))
    assert (cs > 0).all() and (dist > 0).all()
