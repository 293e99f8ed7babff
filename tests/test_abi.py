"""The C ABI library loads without a GPU and exports every function
include/snac_hip.h declares; the ctypes mirror matches the header."""
import ctypes as C
import os
import re

import pytest

from snac_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "snac_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(snh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_step_entry_points():
    names = declared_functions()
    for n in ("snh_context_create", "snh_timestep_zero", "snh_calc_stresses", "snh_loop_nodes_momentum",
              "snh_loop_elements_momentum", "snh_sync", "snh_step", "snh_download_state", "snh_upload_state"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_the_header():
    assert sorted(_abi.SIGNATURES) == declared_functions()


def test_abi_version_and_error_string_without_gpu():
    lib = _abi.load_library()
    assert lib.snh_abi_version() == _abi.ABI_VERSION
    h = C.c_void_p()
    rc = lib.snh_context_create(None, C.byref(h))
    assert rc == _abi.SNH_EINVAL
    assert b"null" in lib.snh_last_error()


def test_struct_sizes_match_c():
    # sizes as laid out by the C compiler for the header's structs
    import subprocess
    import tempfile
    d = tempfile.mkdtemp()
    src = os.path.join(d, "s.c")
    exe = os.path.join(d, "s")
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include "%s"\nint main(){printf("%%zu %%zu %%zu %%zu %%zu %%zu\\n",'
                'sizeof(snh_material),sizeof(snh_velocity_bc),sizeof(snh_config),sizeof(snh_state),'
                'sizeof(snh_scalars),sizeof(snh_kernel_times));'
                'printf("%%zu %%zu\\n",sizeof(snh_temperature_bc),sizeof(snh_thermal));}\n' % HEADER)
    subprocess.check_call(["gcc", "-o", exe, src])
    sizes = list(map(int, subprocess.check_output([exe]).split()))
    assert sizes == [C.sizeof(_abi.Material), C.sizeof(_abi.VelocityBC), C.sizeof(_abi.Config),
                     C.sizeof(_abi.State), C.sizeof(_abi.Scalars), C.sizeof(_abi.KernelTimes),
                     C.sizeof(_abi.TemperatureBC), C.sizeof(_abi.Thermal)]


def test_sync_state_checks_sizes_before_touching_the_gpu():
    """snh_sync_state refuses a mesh beyond its fixed scratch before any HIP
    call, so this runs without a GPU."""
    from snac_amd import problem as P
    lib = _abi.load_library()
    cfg, keep = P.elastic_box_compression(elements=(256, 128, 129)).to_c()   # 4.2 M elements
    rc = lib.snh_sync_state(C.byref(cfg), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), None)
    assert rc == _abi.SNH_EINVAL and b"beyond" in lib.snh_last_error()
