"""The column kernels under the host sanitizers (CPU only).

The host build of the kernels (tests/emu) runs each lane of a workgroup as a
thread and each barrier as a std::barrier, so ThreadSanitizer sees every
LDS exchange: a location one lane writes and another reads with no barrier
ordering the two is a data race, and a report.  AddressSanitizer sees the
LDS block at exactly the size a launch gets (lds_doubles), so an index
outside it, into the scalar area or the int buffer of another column, or
past the scratch, is a report.

Cases: both solvers; ions (the TEC and top ion boundary on scalar slots);
the GPU's forms of the arithmetic (the segmented hydrostatic product);
UseMSIS with the GSWM tides and WACCM with UseDamping (the level-0 values on
the boundary lanes); tiles of 1, 2 and 4 wavefronts (carries in LDS and in
device memory); and the ensemble step's short-column slots, several columns
sharing one LDS block with wavefront-local synchronisation, which must also
give the workgroup run's bits.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from gitm_amd import synth
from tests.helpers import subset

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
SRC = os.path.join(EMU, "sanitize_main.cpp")
MAGIC = 1196840532
ORDER = ("nds", "vvel", "vel", "temp", "rho", "tempunit", "mmm", "gamma", "cmax", "eddy",
         "alt", "grav", "invr", "dalt", "coslat")


def build(kind: str, out_dir: str) -> str:
    exe = os.path.join(out_dir, f"sanitize_{kind}")
    flags = {"thread": ["-fsanitize=thread", "-O1"],
             "address": ["-fsanitize=address", "-fno-omit-frame-pointer", "-O1"]}[kind]
    subprocess.run(["g++", "-std=c++20", "-g", "-ffp-contract=off", "-pthread", *flags,
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "gitm_amd", "csrc"), "-I" + EMU, SRC,
                    "-o", exe], check=True)
    return exe


@pytest.fixture(scope="module")
def exes(tmp_path_factory):
    from concurrent.futures import ThreadPoolExecutor
    d = str(tmp_path_factory.mktemp("sanitize"))
    kinds = ("thread", "address")
    with ThreadPoolExecutor(2) as ex:  # the two builds side by side
        return dict(zip(kinds, ex.map(lambda k: build(k, d), kinds)))


def write_case(path, cols, solver, n_steps, waves, per_wg=0, ions=False, msis=False,
               gswm=False, waccm=False, damping=False, slot_carry_lds=True, device_math=False):
    nc, L = cols["temp"].shape
    flags = (int(ions) | int(msis) << 1 | int(gswm) << 2 | int(waccm) << 3 |
             int(damping) << 4 | int(slot_carry_lds) << 5 | int(device_math) << 6)
    hdr = np.array([MAGIC, nc, L - 4, solver, n_steps, flags, waves, per_wg], dtype=np.int32)
    keys = list(ORDER)
    if ions:
        keys += ["ions", "ivel", "sza", "mlat"]
    if msis:
        keys += ["msis"]
    if gswm or waccm:
        keys += ["tides"]
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        for k in keys:
            f.write(np.ascontiguousarray(cols[k], dtype=np.float64).tobytes())


def run(exe, path, kind):
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66"
    env["ASAN_OPTIONS"] = "halt_on_error=1 exitcode=67 detect_leaks=0"
    r = subprocess.run([exe, path], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (kind, r.returncode,
                                                             r.stderr[-4000:])
    assert "Sanitizer" not in r.stderr, r.stderr[-4000:]


CASES = [
    # name, grid (nAlts, stretched), columns, solver, steps, waves, per_wg, options
    ("ausm_50", (50, False), [0, 40], 1, 2, 1, 0, {}),
    ("rusanov_50_w2", (50, False), [3, 70], 0, 2, 2, 0, {}),
    ("ausm_100_w4", (100, True), [5], 1, 1, 4, 0, {}),
    ("ausm_ions_100", (100, True), [7], 1, 1, 2, 0, {"ions": True}),
    ("rusanov_ions_50_w4", (50, False), [9, 11], 0, 1, 4, 0, {"ions": True}),
    ("ausm_msis_gswm_50", (50, False), [2, 30], 1, 2, 1, 0, {"msis": True, "gswm": True}),
    ("rusanov_waccm_damping_100", (100, True), [4], 0, 1, 2, 0,
     {"waccm": True, "damping": True}),
    ("ausm_slots_20", (20, False), [0, 1, 2, 3, 4], 1, 2, 1, 2, {}),
    ("rusanov_slots_56_device_carries", (56, False), [1, 2, 3], 0, 1, 1, 3,
     {"slot_carry_lds": False}),
    ("ausm_slots_ions_40", (40, False), [5, 6, 7, 8], 1, 1, 1, 4, {"ions": True}),
    # the GPU's forms (segmented hydrostatic product, log-space sums)
    ("ausm_100_w2_device", (100, True), [6], 1, 1, 2, 0, {"device_math": True}),
    ("ausm_10_device", (10, False), [1, 2], 1, 2, 1, 0, {"device_math": True}),
    ("ausm_slots_50_device", (50, False), [3, 4, 5], 1, 1, 1, 3, {"device_math": True}),
]


@pytest.mark.parametrize("kind", ["thread", "address"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_kernel_under_sanitizer(exes, tmp_path, kind, case):
    name, (na, stretched), idx, solver, steps, waves, per_wg, opt = case
    ions = opt.get("ions", False)
    cols = synth.make_columns(synth.GridSpec(1, 1, na, stretched=stretched), seed=na + solver,
                              ions=ions)
    cols = subset(cols, idx)
    if opt.get("msis"):
        cols["msis"] = synth.msis_bc_values(cols, seed=3)
    if opt.get("gswm") or opt.get("waccm"):
        cols["tides"] = synth.tides_values(cols, seed=3, absolute=opt.get("waccm", False))
    path = str(tmp_path / f"{name}.bin")
    write_case(path, cols, solver, steps, waves, per_wg, ions=ions, msis=opt.get("msis", False),
               gswm=opt.get("gswm", False), waccm=opt.get("waccm", False),
               damping=opt.get("damping", False),
               slot_carry_lds=opt.get("slot_carry_lds", True),
               device_math=opt.get("device_math", False))
    run(exes[kind], path, kind)
