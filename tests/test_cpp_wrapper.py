"""The C++ simulation-object wrapper (include/clothmi.hpp, the IPhysicsbody-shaped
caller of the C ABI): it compiles on its own, its driver program is built and
linked against the library, and on the GPU the driver's FixedUpdate / Update loop
gives exactly the positions and normals of the same substeps run through Python."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from licenta_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "tests", "cpp", "wrapper_demo")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "clothmi.hpp"\nint main() { return sizeof(clothmi::ClothBodyHip) > 0 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                    f"-I{os.path.join(ROOT, 'include')}", str(src)], check=True)


def test_demo_built_and_checks_arguments():
    assert os.path.exists(DEMO), "build() builds tests/cpp/wrapper_demo"
    r = subprocess.run([DEMO], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_wrapper_matches_python_body(tmp_path):
    from licenta_amd._abi import make_params
    from licenta_amd.body import ClothBody
    print("cpp wrapper demo on a 24x24 cloth, 10 frames", flush=True)
    sc = configs.config_scene(0, 24)
    w = float(sc["inv_mass"][sc["inv_mass"] > 0][0])
    assert np.all((sc["inv_mass"] == 0) | (sc["inv_mass"] == w))
    prm = make_params(sc["params"])
    (tmp_path / "params.bin").write_bytes(C.string_at(C.addressof(prm), C.sizeof(prm)))
    sc["xyz"].astype(np.float32).tofile(tmp_path / "xyz.f32")
    sc["tri"].astype(np.uint32).tofile(tmp_path / "tri.u32")
    (sc["inv_mass"] == 0).astype(np.uint8).tofile(tmp_path / "fixed.u8")
    np.array([w], np.float32).tofile(tmp_path / "inv_mass.f32")
    r = subprocess.run([DEMO, str(tmp_path), "1", "10"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    x = np.fromfile(tmp_path / "out_xyz.f32", np.float32).reshape(-1, 3)
    nrm = np.fromfile(tmp_path / "out_nrm.f32", np.float32).reshape(-1, 3)
    body = ClothBody(sc["xyz"], sc["tri"], sc["params"], (sc["inv_mass"] == 0).astype(np.uint8),
                     sc["inv_mass"])
    body.fixed_update(10)
    assert np.array_equal(x, body.positions())
    assert np.array_equal(nrm, body.normals())
    body.close()
