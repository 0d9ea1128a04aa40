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


def _c_layout(structs):
    """sizeof and field offsets of header structs, from gcc on the header."""
    import tempfile
    hdr = os.path.join(os.path.dirname(_lib.LIB_PATH), "..", "include", "firewave.h")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.abspath(hdr)}"',
             "int main(void) {"]
    for name, fields in structs.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for fld in fields:
            lines.append(f'  printf("{name} {fld} %zu\\n", offsetof({name}, {fld}));')
    lines += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write("\n".join(lines))
        subprocess.run(["gcc", "-Wall", src, "-o", exe], check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        res[(s, f)] = int(v)
    return res


def test_struct_layouts_match_ctypes():
    """The ctypes mirrors of the header structs have the C sizes and offsets."""
    from firewebapp_amd.physics_3d import lagrangian_bed_3d as LB
    from firewebapp_amd.spread_3d import FwSpreadFields, FwSpreadParams, FwSpreadState
    mirrors = {"fw_grid": _lib.FwGrid, "fw_spread_fields": FwSpreadFields,
               "fw_spread_params": FwSpreadParams, "fw_spread_state": FwSpreadState,
               "fw_bed_particles": LB.FwBedParticles, "fw_bed_index": LB.FwBedIndex,
               "fw_bed_params": LB.FwBedParams}
    fields = {k: [f[0] for f in v._fields_] for k, v in mirrors.items()}
    c = _c_layout(fields)
    for name, cls in mirrors.items():
        assert c[(name, "size")] == ctypes.sizeof(cls), name
        for fld in fields[name]:
            assert c[(name, fld)] == getattr(cls, fld).offset, (name, fld)
