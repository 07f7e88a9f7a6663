"""The C-ABI boundary: the shared library loads without a GPU, exports every
entry point include/samos_hip.h declares, and reports errors through
samos_last_error without a context."""
import ctypes
import subprocess

import pytest

import samos_amd


def test_library_exports_every_declared_symbol():
    lib = samos_amd.load_library()
    names = samos_amd.exported_symbols()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", samos_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(lib, n)


def test_field_codes_match_header():
    text = open(samos_amd.HEADER_PATH).read()
    assert "SAMOS_X = 0" in text and "SAMOS_TYPE = 32" in text
    order = ["SAMOS_X", "SAMOS_Y", "SAMOS_Z", "SAMOS_VX", "SAMOS_VY", "SAMOS_VZ", "SAMOS_NX", "SAMOS_NY",
             "SAMOS_NZ", "SAMOS_FX", "SAMOS_FY", "SAMOS_FZ", "SAMOS_TAUX", "SAMOS_TAUY", "SAMOS_TAUZ",
             "SAMOS_OMEGA", "SAMOS_AGE", "SAMOS_RADIUS", "SAMOS_MASS"]
    pos = [text.index(o + ",") if o + "," in text else text.index(o) for o in order]
    assert pos == sorted(pos)
    assert len(order) == len(samos_amd.DOUBLE_FIELDS)


def test_create_error_is_reported_without_context():
    from tests.conftest import gpu_available
    lib = samos_amd.load_library()
    h = ctypes.c_void_p()
    rc = lib.samos_create(ctypes.byref(h), 0, -5, 10.0, 10.0, 10.0, 1)
    assert rc != 0
    assert lib.samos_last_error(None).decode()
    if not gpu_available():
        with pytest.raises(samos_amd.SamosError):
            samos_amd.System(10, (10.0, 10.0, 10.0))
