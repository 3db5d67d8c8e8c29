"""The C ABI: the library exports what include/gitm_vert.h declares, the
ctypes mirror matches the C struct layout, and the host-only helpers work.
No GPU calls."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from gitm_amd import _abi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gitm_vert.h")
LIB = os.path.join(ROOT, "gitm_amd", "libgitm_vert.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(gitm_(?:vert|horiz)_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("gitm_vert_create", "gitm_vert_destroy", "gitm_vert_advance",
              "gitm_vert_advance_host", "gitm_vert_get_error", "gitm_vert_set_stream",
              "gitm_vert_set_tile", "gitm_vert_default_params", "gitm_vert_strerror",
              "gitm_vert_sources", "gitm_vert_sources_host", "gitm_horiz_advance",
              "gitm_horiz_advance_host", "gitm_horiz_exchange"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgitm_vert.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gitm_(?:vert|horiz)_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgitm_vert.so not built")
def test_host_helpers():
    lib = ctypes.CDLL(LIB)
    lib.gitm_vert_strerror.restype = ctypes.c_char_p
    assert lib.gitm_vert_strerror(0) == b"ok"
    assert b"friction" in lib.gitm_vert_strerror(2)
    p = _abi.Params()
    lib.gitm_vert_default_params(ctypes.byref(p))
    # ModInputs.f90 defaults and init_planet masses
    assert p.beta_limiter == 2.0 and p.max_vertical_velocity == 1000.0
    assert p.omega_body == synth.StepParams().omega_body
    assert list(p.mass) == list(synth.species_masses())
    assert p.thermal_diff_coef[5] == -0.38 and p.use_neutral_friction == 1
    # null handle is rejected without touching a device
    assert lib.gitm_vert_advance(None, ctypes.byref(p), None, 1, 1) == 4


def test_struct_layout_matches_header():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "gitm_vert.h"
    int main(void) {
      printf("%zu %zu %zu %zu\n", sizeof(gitm_vert_field), sizeof(gitm_vert_columns),
             sizeof(gitm_vert_params), sizeof(gitm_vert_error));
      printf("%zu %zu %zu %zu\n", offsetof(gitm_vert_columns, ndensity_s),
             offsetof(gitm_vert_columns, mlat), offsetof(gitm_vert_params, mass),
             offsetof(gitm_vert_params, use_damping));
      return 0;
    }
    """
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "t.c")
        exe = os.path.join(td, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = list(map(int, lines))
    want = [ctypes.sizeof(_abi.Field), ctypes.sizeof(_abi.Columns),
            ctypes.sizeof(_abi.Params), ctypes.sizeof(_abi.Error),
            _abi.Columns.ndensity_s.offset, _abi.Columns.mlat.offset,
            _abi.Params.mass.offset, _abi.Params.use_damping.offset]
    assert got == want
