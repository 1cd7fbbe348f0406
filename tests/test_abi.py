"""The C ABI library exports every entry point include/reapers_hip.h
declares, and the product fails loudly (never falls back to the CPU) when no
GPU is present.  No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "reapers_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("rp_init", "rp_apply_ops", "rp_terms_upload", "rp_gs_pass", "rp_lincomb",
                 "rp_vec_prod", "rp_vec_norm", "rp_copy_vec", "rp_axpy", "rp_scale_vec"):
        assert must in names


def test_library_exports_every_declared_symbol(ra):
    lib = ctypes.CDLL(ra.HIP_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_cpu_library_exports_every_declared_symbol(ra):
    """include/reapers_cpu.h (the REAPERS_NOGPU build of the C++ API) is
    exported by libreapers_cpu.so, which links no HIP library."""
    from reapers_amd import cpu
    text = open(os.path.join(ROOT, "include", "reapers_cpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(rpc_[a-z0-9_]+)\s*\(", text)))
    assert "rpc_evolve" in names and "rpc_apply" in names
    lib = ctypes.CDLL(cpu.CPU_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    import subprocess
    deps = subprocess.run(["ldd", cpu.CPU_LIB], capture_output=True, text=True).stdout
    assert "amdhip" not in deps and "reapers_hip" not in deps


def test_host_library_loads(ra):
    ra.host()


def test_calls_before_init_fail_cleanly(ra):
    h = ra.hip()
    p = ctypes.c_void_p()
    if h.rp_malloc(ctypes.byref(p), 16) == 0:
        pytest.skip("the backend is already initialized in this process")
    assert h.rp_malloc(ctypes.byref(p), 16) == 5          # RP_ERR_NOT_INITIALIZED
    assert b"rp_init" in h.rp_last_error()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_no_gpu_fails_loudly(ra):
    """Without a GPU the backend reports RP_ERR_NO_DEVICE, and the C++ API
    throws NoGpuDevice instead of computing anything on the host."""
    assert ra.hip().rp_init(0) == 4
    h = ra.host().rph_state_new(32, 4)
    assert not h
    assert b"No GPU device available" in ra.host().rph_last_error()


def test_cpp_api_argument_errors(ra):
    """Errors the reference throws before touching the device."""
    host = ra.host()
    assert not host.rph_state_new(32, 0)
    assert b"len" in host.rph_last_error() and b"cannot be zero" in host.rph_last_error()
    assert not host.rph_state_new(64, 33)
    bits = (ctypes.c_uint64 * 1)()
    co = (ctypes.c_double * 2)()
    assert host.rph_majorana(64, 8, 9, 1, bits, co) == 1     # FieldIndexTooLarge
    assert b"field index 9 exceeds 8" in host.rph_last_error()
    assert host.rph_majorana(64, 7, 1, 1, bits, co) == 1     # N must be even
    n = ctypes.c_int64()
    assert host.rph_syk_gen(64, 3, 0.0, 1.0, 1, 0, 0, bits, co, 1, ctypes.byref(n)) == 1
