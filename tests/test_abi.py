"""The C-ABI library loads and exports every symbol include/ljgpu.h declares;
host-only entry points validate their arguments. No device calls here."""
import ctypes as C
import re
import subprocess

import pytest

import lj_gpu_amd as G


def test_library_loads_and_exports_every_declared_symbol():
    L = G.lib()
    names = G.header_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    # and the shared object's dynamic symbol table agrees
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (ljgpu_\w+)", out))
    assert set(names) <= exported
    # nothing of the oracle is linked into the product
    assert "orc_" not in out


def test_library_targets_gfx950_only():
    data = open(G.LIB_PATH, "rb").read()
    targets = set(re.findall(rb"amdgcn-amd-amdhsa--(gfx[0-9a-z]+)", data))
    assert targets == {b"gfx950"}


def test_version_string():
    assert b"gfx950" in G.lib().ljgpu_version()


def test_error_codes_without_arguments():
    L = G.lib()
    assert L.ljgpu_create(0, None) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_device_count(None) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_sync(None) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_destroy(None) == G.OK
    assert L.ljgpu_force_aos_f64(None, None, None, 0.01, 9.0, 0.0, 1) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_set_list_csr(None, 0, None, None, None, 0, 0, 0) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_lattice_reference(-1.0, 50.0, None, 0, None) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_validate_csr(-1, None, None, None, 0, 0) == G.ERR_INVALID_ARGUMENT
    assert L.ljgpu_validate_csr(0, None, None, None, 0, 7) == G.ERR_INVALID_ARGUMENT
    assert G.lib().ljgpu_last_error()


def test_no_device_is_an_error_not_a_crash():
    if G.device_count() > 0:
        pytest.skip("a device is present")
    h = C.c_void_p()
    assert G.lib().ljgpu_create(0, C.byref(h)) == G.ERR_NO_DEVICE
    assert h.value is None
    with pytest.raises(G.LJGPUError):
        G.Context(0)


def test_buffer_size_checks_reject_short_buffers():
    """The Python mirror checks device buffer sizes before calling into the
    library (the C ABI has bare pointers); host-side only."""
    class Fake:
        def __init__(self, nbytes, ptr=1):
            self.nbytes, self.ptr = nbytes, ptr
    G._need(Fake(32 * 10), 32 * 10, "q")
    with pytest.raises(ValueError):
        G._need(Fake(32 * 10 - 8), 32 * 10, "q")
    with pytest.raises(ValueError):
        G._need(Fake(64, ptr=None), 32, "q")
    with pytest.raises(ValueError):
        G._need(None, 0, "q")
