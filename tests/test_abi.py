"""CPU: the C-ABI library loads and exports every function include/nyxpm.h
declares; the Python binding covers the same set.  No compute calls (no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nyxpm.h")
LIB = os.path.join(ROOT, "nyx_amd", "libnyxpm.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nyxpm_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ["nyxpm_assign_density", "nyxpm_solve_for_phi", "nyxpm_get_grav_vector",
                 "nyxpm_move_kick_drift", "nyxpm_move_kick", "nyxpm_pm_steps", "nyxpm_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libnyxpm.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (nyxpm_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for n in declared():
        getattr(lib, n)


def test_binding_matches_header():
    from nyx_amd import _lib
    bound = sorted(n for n, _, _ in _lib.SIGNATURES)
    assert bound == declared()


def test_null_context_is_an_error_not_a_crash():
    from nyx_amd import _lib
    L = _lib.load()
    assert L.nyxpm_sync(None) == _lib.NYXPM_ERR_ARG
    assert L.nyxpm_destroy(None) == 0
    assert L.nyxpm_last_error(None) == b"null context"
    assert L.nyxpm_num_particles(None, 0) == -1
