"""The C-ABI library loads and exports every entry point include/pyrecon_amd.h declares.

No compute call is made (this runs without a GPU).  The signatures registered in
the ctypes binding (pyrecon_amd/_lib.py) must cover exactly the declared set.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'pyrecon_amd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(prh_\w+)\s*\(', text)))


def test_header_parses():
    names = declared_functions()
    assert len(names) > 40
    for name in ['prh_assign_cic_f64', 'prh_read_finite_difference_cic_f32', 'prh_fmg_f64', 'prh_fmg_mixed',
                 'prh_paint_f64', 'prh_read_shifts_f64', 'prh_smooth_gaussian_f64', 'prh_workspace_create']:
        assert name in names


def test_library_exports_every_declared_symbol():
    from pyrecon_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_binding_covers_header():
    from pyrecon_amd import _lib
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_library_is_in_tree_and_for_gfx950():
    from pyrecon_amd import _lib
    assert os.path.dirname(_lib.LIB_PATH) == os.path.join(ROOT, 'pyrecon_amd')
    # the code object of the library targets gfx950 (the offload bundle names its target)
    data = open(_lib.LIB_PATH, 'rb').read()
    assert b'gfx950' in data


def test_error_paths_without_device():
    # argument validation happens before any device work
    from pyrecon_amd import _lib
    lib = _lib.lib
    assert lib.prh_workspace_create(None, None, 8, 1, 1) == _lib.ERR_ARG
    nm, nm_ptr = _lib.size3([1, 8, 8])
    handle = ctypes.c_void_p()
    assert lib.prh_workspace_create(ctypes.byref(handle), nm_ptr, 8, 16, _lib.WS_MULTIGRID) == _lib.ERR_ARG
    assert 'out of the supported range' in _lib.last_error()
    nm, nm_ptr = _lib.size3([8, 8, 8])
    assert lib.prh_workspace_create(ctypes.byref(handle), nm_ptr, 3, 16, _lib.WS_MULTIGRID) == _lib.ERR_ARG
    assert lib.prh_timer_enable(99, 1) == _lib.ERR_ARG
    with pytest.raises(ValueError):
        _lib.check(_lib.ERR_ARG, 'x')
    with pytest.raises(_lib.MeshError):
        _lib.check(_lib.ERR_RANGE, 'x')
