"""The C ABI library loads without a GPU, exports every function that
include/firex.h declares, and fails cleanly (no partial state, a FIREX_ERR_*
status and a message) where the device is needed but absent or where the
request breaks a host-side rule."""
import ctypes
import os
import re
import subprocess

import pytest

from fire_amd import solver as fs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", n) for n in sorted(os.listdir(os.path.join(ROOT, "include")))
           if n.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(firex_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", fs.LIBRARY_PATH], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


def test_header_declares_the_solver_api():
    names = declared_functions()
    for required in ("firex_create", "firex_destroy", "firex_step", "firex_read_field", "firex_write_field",
                     "firex_reduce", "firex_set_parameter", "firex_render_offline_frame", "firex_frame"):
        assert required in names


def test_library_exports_every_declared_function():
    exported = exported_symbols()
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared in include/ but not exported: {missing}"


def test_python_mirror_binds_every_export():
    from fire_amd import serve

    lib = fs.load_library()
    signatures = dict(fs._SIGNATURES, **serve.SERVE_SIGNATURES)
    for name in declared_functions():
        assert hasattr(lib, name)
        assert name in signatures, f"{name} has no ctypes signature in fire_amd/solver.py or fire_amd/serve.py"


def test_abi_version():
    assert fs.abi_version() == 1


def test_budget_overflow_is_rejected_before_any_allocation():
    with pytest.raises(fs.FireXError) as e:
        fs.FireXHipSolver((192, 288, 192), tier="uhd", memory_budget_bytes=64 * 1024 * 1024)
    assert e.value.code == fs.ERR_BUDGET
    assert "guard" in str(e.value)


@pytest.mark.parametrize("kwargs,code", [
    (dict(dims=(1, 24, 16)), fs.ERR_UNSUPPORTED),
    (dict(dims=(16, 24, 2000)), fs.ERR_UNSUPPORTED),
    (dict(dims=(16, 24, 16), pressure_iterations=-1), fs.ERR_INVALID),
    (dict(dims=(16, 24, 16), time_step=0.0), fs.ERR_INVALID),
    (dict(dims=(16, 24, 16), maximum_substeps=0), fs.ERR_INVALID),
])
def test_invalid_configurations(kwargs, code):
    with pytest.raises(fs.FireXError) as e:
        fs.FireXHipSolver(**kwargs)
    assert e.value.code == code


def test_create_without_a_device_fails_loudly():
    if fs.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(fs.FireXError) as e:
        fs.FireXHipSolver((16, 24, 16))
    assert e.value.code == fs.ERR_DEVICE
    assert "No HIP device" in str(e.value)


def test_null_arguments_are_errors_not_crashes():
    lib = fs.load_library()
    assert lib.firex_create(None, None) == fs.ERR_INVALID
    assert lib.firex_step(None, 0.1, 1) == fs.ERR_INVALID
    assert lib.firex_sync(None) == fs.ERR_INVALID
    assert lib.firex_read_field(None, 0, None, 0, 0) == fs.ERR_INVALID
    lib.firex_destroy(None)
    out = ctypes.c_double()
    assert lib.firex_clamp_parameter(None, 1.0, 1.0, ctypes.byref(out)) == fs.ERR_INVALID


# ctypes mirror of each C struct: (ctypes class, C struct name)
def _struct_pairs():
    from fire_amd import serve as sv
    return [
        (fs.Config, "firex_config"), (fs.Emitter, "firex_emitter"), (fs.Info, "firex_info"),
        (fs.MemoryEstimate, "firex_memory_estimate"), (fs.DeviceLimits, "firex_device_limits"),
        (fs.PassTiming, "firex_pass_timing"), (sv.ServerConfig, "firex_server_config"),
        (sv.Request, "firex_request"), (sv.RequestStatus, "firex_request_status"),
        (sv.ServerStats, "firex_server_stats"), (fs.FieldLayout, "firex_field_layout"),
        (fs.RenderConfig, "firex_render_config"),
    ]


def test_python_structs_match_the_c_layout(tmp_path):
    """Every ctypes mirror has the C struct's size and field offsets (gcc
    compiles the headers and prints sizeof / offsetof)."""
    include = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "firex.h"', '#include "firex_serve.h"',
             "int main(void) {"]
    expect = []
    for cls, cname in _struct_pairs():
        lines.append(f'  printf("%zu\\n", sizeof({cname}));')
        expect.append((f"sizeof {cname}", ctypes.sizeof(cls)))
        for fname, _ in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append((f"{cname}.{fname}", getattr(cls, fname).offset))
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", include, str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert len(got) == len(expect)
    bad = [(name, c, py) for (name, py), c in zip(expect, got) if c != py]
    assert not bad, f"C / ctypes layout mismatches: {bad}"
