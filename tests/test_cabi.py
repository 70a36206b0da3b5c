"""The C-ABI library loads without a GPU and exports every function that
include/smmp_mi355x.h declares (no compute calls)."""
import ctypes
import os
import re
import subprocess

from conftest import REPO


def declared_functions():
    with open(os.path.join(REPO, "include", "smmp_mi355x.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)  # comments may show call syntax
    return sorted(set(re.findall(r"\b(smmp_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("smmp_energy", "smmp_metropolis", "smmp_partem_s", "smmp_energy_batch", "smmp_sample",
                 "smmp_batch_create", "smmp_batch_run"):
        assert must in names


def test_library_exports_every_declared_function():
    from smmp_amd import _ffi
    lib = ctypes.CDLL(_ffi.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and as dynamic symbols of the shared object itself
    nm = subprocess.run(["nm", "-D", "--defined-only", _ffi.LIBPATH], capture_output=True, text=True, check=True)
    syms = {ln.split()[-1] for ln in nm.stdout.splitlines() if ln.strip()}
    assert not [n for n in declared_functions() if n not in syms]


def test_binding_covers_the_header():
    from smmp_amd import _ffi
    assert set(_ffi.EXPORTS) == set(declared_functions())
