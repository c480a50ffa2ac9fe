"""CPU checks of the drop-in boundary: libempm.so loads, exports every symbol
include/empm.h declares, and the ctypes mirrors of its structs have the C
layout (offsets compiled from the header by gcc).  No compute call needs a
GPU here; argument errors are rejected before any HIP call."""
import ctypes as C
import os
import re
import subprocess

import pytest

from empm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "empm.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(empm_[a-z_0-9]+)\s*\(", src)))


def test_library_loads_and_abi_version():
    lib = _lib.load()
    assert lib.empm_abi_version() == _lib.EMPM_ABI_VERSION


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    declared = header_functions()
    assert len(declared) >= 10
    assert sorted(declared) == sorted(_lib.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (empm_\w+)", out))
    assert set(declared) <= exported, set(declared) - exported


STRUCTS = {
    "empm_create_desc": _lib.CreateDesc,
    "empm_model": _lib.Model,
    "empm_particles": _lib.Particles,
    "empm_collider": _lib.Collider,
    "empm_grid_rec": _lib.GridRec,
    "empm_batch": _lib.Batch,
    "empm_batch_create_desc": _lib.BatchCreateDesc,
}


def test_ctypes_layouts_match_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            k, v = ln.rsplit(" ", 1)
            got[k] = int(v)
    for cname, py in STRUCTS.items():
        assert got[f"{cname} size"] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[f"{cname}.{f}"] == getattr(py, f).offset, (cname, f)


def test_argument_errors_are_reported_without_gpu():
    lib = _lib.load()
    h = C.c_void_p()
    desc = _lib.CreateDesc(device=0, n_grid=2, max_particles=10)
    rc = lib.empm_create(C.byref(h), C.byref(desc))
    assert rc == _lib.EMPM_E_INVALID_ARG
    assert b"n_grid" in lib.empm_last_error()
    assert lib.empm_mu_lam(-1, None, None, None, None, None) == _lib.EMPM_E_INVALID_ARG
    with pytest.raises(_lib.EmpmError):
        _lib.check(lib.empm_substep_fwd(None, None, None, 0, C.c_float(0), C.c_float(1e-4), None, None, None))


def test_product_path_has_no_oracle_or_cpu_fallback():
    """The package must not import the oracle, and must fail loudly when the
    HIP library is missing."""
    pkg = os.path.join(ROOT, "empm_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith(".py"):
                txt = open(os.path.join(dp, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", txt, flags=re.M), f
    code = ("import empm_amd._lib as L\n"
            "try:\n    L.load('/nonexistent/libempm.so')\nexcept ImportError as e:\n    print('raised')\n")
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, cwd=ROOT)
    assert "raised" in r.stdout
