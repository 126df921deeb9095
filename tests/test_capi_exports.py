"""CPU checks of the C-ABI library: it builds for gfx950, loads, and exports
every function include/*.h declares (no device calls: there is no GPU here)."""
import ctypes
import glob
import os
import re
import shutil
import subprocess

import pytest

import z4c_amd
from z4c_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(z4c_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exists_and_loads():
    assert os.path.exists(_capi.LIB_PATH), "build the library first (__graft_entry__.build())"
    L = _capi.load()
    assert L.z4c_abi_version() == 1


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 19
    L = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_capi.EXPORTS) == names


def test_no_torch_types_in_header():
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        assert "torch" not in src.lower().replace("torch.autograd", "")
        assert "hip_runtime" not in src  # plain pointers and sizes only


def test_binary_is_gfx950(tmp_path):
    # --offloading extracts the device images next to its input: work on a copy
    lib = tmp_path / "lib.so"
    shutil.copy(_capi.LIB_PATH, lib)
    out = subprocess.run(["/opt/rocm/llvm/bin/llvm-objdump", "--offloading", str(lib)],
                         capture_output=True, text=True, cwd=tmp_path)
    if out.returncode != 0:
        pytest.skip("llvm-objdump --offloading unavailable")
    assert "gfx950" in out.stdout + out.stderr


def test_params_layout_matches_header():
    # 17 doubles + 4 ints
    assert ctypes.sizeof(_capi.Params) == 17 * 8 + 4 * 4
    assert ctypes.sizeof(_capi.Block) == 8 + 9 * 4 + 4 + 6 * 8 + 4 * 4


def test_low_storage_tableaus_equal_the_oracle():
    """z4c_ls_tableau_get (host only, no device) returns the reference's
    LowStorage weights, the ones the oracle's integrator tests pin; the
    integrators GR_Z4c rejects or does not know give Z4C_EINVAL."""
    import oracle as orc
    L = _capi.load()
    for name, stages in orc.LS_TABLEAUS.items():
        t = _capi.LsTableau()
        assert L.z4c_ls_tableau_get(name.encode(), ctypes.byref(t)) == 0
        assert t.nstages == len(stages)
        assert t.stages() == [tuple(float(v) for v in st) for st in stages], name
    t = _capi.LsTableau()
    for bad in (b"ssprk5_4", b"bt_rk4", b""):
        assert L.z4c_ls_tableau_get(bad, ctypes.byref(t)) == -1
    assert ctypes.sizeof(_capi.LsTableau) == 8 + 5 * 5 * 8
