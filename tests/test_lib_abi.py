"""The C-ABI library loads and exports every symbol include/phf_hip.h declares
(no GPU needed: no compute calls)."""
import os
import re

from phf_pcc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(phf_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    assert {"phf_buildhsg", "phf_localg", "phf_set_hamiltonian"} <= _declared()


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in sorted(_declared()):
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) == _declared()


def test_last_error_without_context_is_callable():
    assert isinstance(_lib.load().phf_last_error(None), bytes)


def test_sc_entry_points_reject_null_context():
    """Argument marshalling of the BuildSC entry points (no GPU: a NULL context
    returns PHF_ERR_ARG before any device call)."""
    import numpy as np
    lib = _lib.load()
    z = np.zeros(64, np.complex128)
    ng = np.array([1, 1], np.int32)
    ip = ng.ctypes.data_as(_lib._ip)
    p = _lib.ptr(z.view(np.float64))
    assert lib.phf_sc_set_grid(None, 4, 2, ip, 0, p, 1, p, 1, p, 1, p, p, 1, p, 1, p) == _lib.PHF_ERR_ARG
    assert lib.phf_sc_apply(None, 1, p, p, p, p, p, p) == _lib.PHF_ERR_ARG
    assert lib.phf_sc_apply_packed(None, 1, p, p) == _lib.PHF_ERR_ARG
    assert lib.phf_sc_apply_packed_dev(None, 1, None, None) == _lib.PHF_ERR_ARG
    assert lib.phf_buildsc(None, p, p, p, 2, 4, ip, 0, p, 1, p, 1, p, 1, p, p, 1, p, 1, p, p, p, p) == _lib.PHF_ERR_ARG
    assert lib.phf_sd_dim(8, 3) == 1 + 15 + 3 * 10
    assert lib.phf_sc_set_points(None, 4, 1, None, p, None) == _lib.PHF_ERR_ARG
    assert lib.phf_eval_objective(None, p, 1, 2, 0, p, p) == _lib.PHF_ERR_ARG
    info = np.zeros(1, np.int32)
    assert lib.phf_invert_batched(None, 4, 1, p, p, p, info.ctypes.data_as(_lib._ip)) == _lib.PHF_ERR_ARG


def test_buildsc_shim_checks_shapes_before_the_device():
    import numpy as np
    import pytest
    from phf_pcc_amd.PHFTools_F import phftools
    n = 4
    eye = np.eye(n, dtype=complex)[None]
    args = (eye, eye, eye, eye, np.zeros(1), np.zeros(1), np.zeros(1), np.ones((1, 1)), np.ones((1, 1, 1)),
            np.ones(1, complex), np.ones(1, complex), np.ones(1, complex))
    with pytest.raises(ValueError, match="c2 must be"):
        phftools.buildsc(1.0, np.zeros((n, n)), np.zeros((n, n, n)), 2, [1, 1], 0, *args)
    with pytest.raises(TypeError, match="missing required argument 'fk'"):
        phftools.buildsc(1.0, np.zeros((n, n)), np.zeros((n,) * 4), 2, [1, 1], 0, 1, *args[:-1])
