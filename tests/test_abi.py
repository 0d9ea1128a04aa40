"""CPU-side checks of the C ABI: the library loads and exports every symbol
the header declares; the header parses into ctypes signatures."""
import ctypes
import os
import subprocess

import pytest

from firewebapp_amd import _lib


def test_header_parses():
    decls = _lib.parse_header()
    assert "fw_last_error" in decls and "fw_species_transport" in decls
    for name, (res, args) in decls.items():
        assert name.startswith("fw_")


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfirewave.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in _lib.parse_header() if n not in exported]
    assert not missing, missing
    L = _lib.lib()  # dlopen + argtypes for every declaration
    assert L.fw_version() >= 1


def test_config_error_path_needs_no_gpu():
    """A null grid is rejected on the host with FW_E_CONFIG."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfirewave.so not built")
    L = _lib.lib()
    rc = L.fw_eos(None, None, None, 300.0, None)
    assert rc == _lib.FW_E_CONFIG
    assert b"grid" in L.fw_last_error()
    with pytest.raises(_lib.FirewaveError):
        _lib.check(rc, "fw_eos")
