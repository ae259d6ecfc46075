"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, the ctypes descriptor matches the C struct layout, and the
host-only entry points behave.  No device calls."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from taichirtm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "taichirtm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(trtm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_abi_version():
    assert _lib.load().trtm_abi_version() == _lib.ABI_VERSION == 7


@pytest.mark.parametrize("cname,pystruct", [("trtm_step_desc", _lib.StepDesc),
                                            ("trtm_fwi_desc", _lib.FwiDesc)])
def test_descriptor_layout_matches_c(cname, pystruct):
    fields = [f for f, _ in pystruct._fields_]
    prog = ['#include <stdio.h>', '#include <stddef.h>', '#include "taichirtm_hip.h"',
            "int main(void) {", '  printf("%%zu\\n", sizeof(%s));' % cname]
    prog += ['  printf("%%zu\\n", offsetof(%s, %s));' % (cname, f) for f in fields]
    prog += ["  return 0;", "}"]
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "layout.c")
        exe = os.path.join(td, "layout")
        open(src, "w").write("\n".join(prog))
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                              check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(pystruct)
    for f, off in zip(fields, out[1:]):
        assert getattr(pystruct, f).offset == off, f


@pytest.mark.parametrize("fw", [0, 1, 20, 40, 100])
def test_sponge_table_matches_oracle(fw):
    tab = np.zeros(fw + 1, np.float32)
    assert _lib.load().trtm_sponge_table(fw, ctypes.c_float(0.0053),
                                         tab.ctypes.data_as(ctypes.c_void_p)) == 0
    n = max(fw, 1) * 2 + 3
    c = oracle.sponge(n, n, fw)
    for d in range(fw):  # left strip value at distance d is table[fw - d]
        assert c[d, 0] == tab[fw - d]


def _plan(nx, nz, loc):
    loc = np.ascontiguousarray(loc, np.int32)
    n = len(loc)
    order = np.zeros(max(n, 1), np.int32)
    start = np.zeros(n + 1, np.int32)
    ng = ctypes.c_int32()
    rc = _lib.load().trtm_plan_injection(nx, nz, n, loc.ctypes.data_as(ctypes.c_void_p),
                                         order.ctypes.data_as(ctypes.c_void_p),
                                         start.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ng))
    return rc, order[:n], start[: ng.value + 1], ng.value


def test_plan_injection_groups_duplicates_in_entry_order():
    loc = [(3, 1), (0, 1), (3, 1), (2, 5), (3, 1), (0, 1)]
    rc, order, start, ng = _plan(8, 8, loc)
    assert rc == 0 and ng == 3
    groups = [list(order[start[g]:start[g + 1]]) for g in range(ng)]
    assert groups == [[1, 5], [3], [0, 2, 4]]  # sorted by cell, stable inside a cell


def test_plan_injection_rejects_outside_grid():
    rc, *_ = _plan(8, 8, [(1, 1), (8, 2)])
    assert rc == _lib.ERR_ARG
    assert b"outside" in _lib.load().trtm_last_error()


def test_plan_injection_empty():
    rc, order, start, ng = _plan(8, 8, np.zeros((0, 2)))
    assert rc == 0 and ng == 0 and list(start) == [0]


def test_step_entry_rejects_bad_descriptor_without_device():
    d = _lib.StepDesc()
    d.struct_size = 7
    rc = _lib.load().trtm_forward_steps(ctypes.byref(d), 0, 1, 0, None)
    assert rc == _lib.ERR_ARG
    assert b"struct_size" in _lib.load().trtm_last_error()
    d.struct_size = ctypes.sizeof(_lib.StepDesc)
    d.nx, d.nz, d.nt, d.fw = 10, 10, 5, 11
    assert _lib.load().trtm_forward_steps(ctypes.byref(d), 0, 1, 0, None) == _lib.ERR_ARG
    assert b"absorbing_frame" in _lib.load().trtm_last_error()


def test_propagator_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    import taichirtm_amd
    from taichirtm_amd import configs
    cfg = configs.psv_anomaly(nx=24, nz=20, nt=10, fw=4, radius=3, src=(12, 5), rec_range=(2, 20))
    kw = cfg.kwargs(cfg.shots[0])
    wu, wv, ww = cfg.wavelets(cfg.shots[0])
    with pytest.raises(RuntimeError, match="HIP device"):
        taichirtm_amd.ForwardModeling(wavelet_u_field=wu, wavelet_v_field=wv, wavelet_w_field=ww,
                                      **kw)


@pytest.mark.parametrize("systems", [_lib.PSV, _lib.SH, _lib.PSV | _lib.SH])
def test_block_configs_enumerate_both_families(systems):
    """The autotuner's search space comes from the library: every compiled
    shape of each temporally blocked family, wave shapes with 64 - 2K columns."""
    from taichirtm_amd import autotune
    shapes = autotune.candidates(systems)
    fams = {s[0] for s in shapes}
    assert fams == {"blocked", "wave"}
    for fam, tx, tz, k in shapes:
        assert tx > 0 and tz > 0 and k > 0
        if fam == "wave":
            assert tz == 64 - 2 * k
    buf = (ctypes.c_int32 * 4)()
    assert _lib.load().trtm_block_configs(systems, 99, buf, 1) < 0  # not a blocked family


def test_plan_entries_check_arguments_without_device():
    """The captured-stepping entries reject missing handles before any GPU work."""
    lib = _lib.load()
    assert lib.trtm_plan_destroy(None) == 0
    assert lib.trtm_plan_launch(None, None) < 0
    assert b"plan is NULL" in lib.trtm_last_error()
    assert lib.trtm_plan_create(None, 0, None, 0, 0, 1, 1, 0, None, None) < 0
    assert b"plan is NULL" in lib.trtm_last_error()
    handle = ctypes.c_void_p()
    assert lib.trtm_plan_create(None, 0, None, 0, 0, 1, 1, 0, None, ctypes.byref(handle)) < 0
    assert b"empty batch" in lib.trtm_last_error()
    assert handle.value is None
