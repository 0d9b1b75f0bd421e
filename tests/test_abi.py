"""The C-ABI library loads without a GPU and exports every function that
include/clothmi.h declares; the ctypes mirror matches the header's layout."""
import ctypes as C
import os
import re

import numpy as np

from licenta_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "clothmi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_abi.LIB_PATH):
        from licenta_amd import build
        build.build()
    lib = C.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_abi.EXPORTS)


def test_struct_layouts():
    assert C.sizeof(_abi.CmParams) == 144
    assert C.sizeof(_abi.CmContact) == 48
    p = _abi.make_params(__import__("licenta_amd.configs", fromlist=["x"]).default_params())
    assert abs(p.gravity[1] + 0.0981) < 1e-7 and p.morton_bits == 63 and p.pair_cap == 64


def test_default_params_match_library():
    lib = _abi.lib()
    p = _abi.CmParams()
    lib.cm_default_params(C.byref(p))
    q = _abi.make_params(__import__("licenta_amd.configs", fromlist=["x"]).default_params())
    assert bytes(p) == bytes(q)


def test_create_fails_cleanly_on_bad_input():
    lib = _abi.lib()
    p = _abi.CmParams()
    lib.cm_default_params(C.byref(p))
    h = C.c_void_p()
    xyz = np.zeros((3, 3), np.float32)
    tri = np.array([[0, 1, 5]], np.uint32)  # index out of range
    inv = np.ones(3, np.float32)
    rc = lib.cm_create(C.byref(p), _abi.ptr(xyz), 3, _abi.ptr(tri), 1, None, _abi.ptr(inv), None, 0,
                       None, None, C.byref(h))
    assert rc == -1 and h.value is None
    assert b"out of range" in lib.cm_last_error()


def test_entry_points_reject_null_body():
    """Every body entry point checks its handle before touching a device."""
    lib = _abi.lib()
    buf = np.zeros(12, np.float32)
    picks = np.zeros(3, np.int32)
    assert lib.cm_step(None, 1) == -1
    assert lib.cm_set_stiffness(None, None, None) == -1
    assert lib.cm_forward_ckpt(None, 4, 2) == -1
    assert lib.cm_adjoint(None, _abi.ptr(buf), None, None, None, None, None) == -1
    assert lib.cm_autotune(None, 2, _abi.ptr(picks)) == -1
    assert b"null body" in lib.cm_last_error()
    assert lib.cm_read_vertices(None, _abi.ptr(buf)) == -1


def test_batch_rejects_per_body_stages_before_any_device_call():
    """cm_create_batch refuses LRA, rigid-mode damping and friction from the
    parameters alone (no device needed), with a message naming them."""
    from licenta_amd import configs
    lib = _abi.lib()
    xyz = np.zeros((3, 3), np.float32)
    xyz[1, 0] = xyz[2, 2] = 1.0
    tri = np.array([0, 1, 2], np.uint32)
    inv = np.ones(3, np.float32)
    n = np.array([3], np.int32)
    t = np.array([1], np.int32)
    for flag in (configs.F_LRA, configs.F_DAMPING, configs.F_FRICTION):
        prm = configs.default_params()
        prm.update(morton_bits=30, morton_normalize=1,
                   flags=configs.F_SELF_COLLISION | flag)
        p = _abi.make_params(prm)
        h = C.c_void_p()
        rc = lib.cm_create_batch(C.byref(p), 1, _abi.ptr(n), _abi.ptr(t), None, _abi.ptr(xyz),
                                 _abi.ptr(tri), None, _abi.ptr(inv), None, None, None, C.byref(h))
        assert rc == -1 and h.value is None
        assert b"friction" in lib.cm_last_error()


def test_flag_and_layout_constants_match_header():
    """configs.F_* / RC_* and the VertexFormat offsets agree with clothmi.h."""
    from licenta_amd import configs
    text = open(os.path.join(ROOT, "include", "clothmi.h")).read()
    defs = dict(re.findall(r"#define (CM_(?:F|RC)_[A-Z0-9_]+) \(1u << (\d+)\)", text))
    assert len(defs) >= 13
    for name, bit in defs.items():
        py = name[3:]
        assert getattr(configs, py) == 1 << int(bit), name
    assert configs.RC_ALL == sum(1 << int(b) for n, b in defs.items() if n.startswith("CM_RC_"))
    lay = dict(re.findall(r"#define CM_VERTEX_([A-Z]+) (\d+)", text))
    assert _abi.VERTEX_DTYPE.itemsize == 4 * int(lay["FLOATS"]) == 48
    assert _abi.VERTEX_DTYPE.fields["position"][1] == 4 * int(lay["POSITION"])
    assert _abi.VERTEX_DTYPE.fields["normal"][1] == 4 * int(lay["NORMAL"])


def test_integration_snippets_compile(tmp_path):
    """The C++ blocks of INTEGRATION.md sections 2-5 (pair query on device buffers,
    the batch, the adjoint, the dead paths) compile against include/clothmi.h as
    written."""
    import shutil
    import subprocess
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    body = text[text.index("## 2."):text.index("## 6.")]
    blocks = re.findall(r"```cpp\n(.*?)```", body, flags=re.S)
    assert len(blocks) == 4
    src = tmp_path / "snippets.cpp"
    src.write_text("\n".join(blocks))
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    r = subprocess.run([hipcc, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-command-line-argument",
                        "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
