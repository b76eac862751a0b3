"""The C-ABI library loads and exports every symbol include/ff_abi.h
declares (no GPU call is made), and the Python binding covers them all."""
import ctypes
import os
import re

import fastfusion_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ff_abi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ff_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    syms = declared_symbols()
    assert len(syms) >= 20
    lib = ctypes.CDLL(fastfusion_amd.lib_path())
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_abi():
    assert set(declared_symbols()) == set(fastfusion_amd.ABI)


def test_library_has_gfx950_code():
    data = open(fastfusion_amd.lib_path(), "rb").read()
    assert b"gfx950" in data


def test_cpp_adapter_compiles_and_links(tmp_path):
    """The header-only C++ adapter (the FusionMipMapCPU-named class) compiles
    and links against libffgpu.so; nothing is executed."""
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "ff_fusion_mipmap_gpu.hpp"\n'
        "int main(int argc, char **) {\n"
        "  if (argc > 5) { ff::FusionMipMapGPU f(0, 0, 0, 0.005f, 0.015f);\n"
        "    ff::DepthView d{nullptr, 640, 480, 1280}; ff::ColorView c{nullptr, 640, 480, 1920};\n"
        "    ff::CameraPose p{}; f.addMap(d, p, c, 1.0f / 5000.0f, 4.0f); f.updateMeshes();\n"
        "    return (int)f.getMeshInterleavedMarchingCubes().faces.size(); }\n"
        "  return 0; }\n")
    lib = os.path.dirname(fastfusion_amd.lib_path())
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", lib,
                           "-lffgpu", "-o", str(tmp_path / "use")])
