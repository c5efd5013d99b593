"""The C ABI: libflipmi.so loads without a GPU and exports every entry point
include/flipmi.h declares; the Python binding names the same set; the
product fails loudly (no CPU fallback) where no GPU is present."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flipmi.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s*(flipmi_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib_path():
    import flipengine3d_amd as fa
    if not os.path.exists(fa.LIB_PATH):
        from flipengine3d_amd import build
        build.build()
    return fa.LIB_PATH


def test_header_declares_the_entry_points():
    names = declared()
    assert len(names) >= 20
    for n in ("flipmi_scene_update", "flipmi_pressure_solve", "flipmi_p2g", "flipmi_advect_particles"):
        assert n in names


def test_library_exports_every_declared_symbol(lib_path):
    lib = ctypes.CDLL(lib_path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    import flipengine3d_amd as fa
    assert sorted(fa.EXPORTED) == declared()


def test_struct_layouts_match_header():
    """The ctypes mirrors have the sizes the C compiler gives the header's structs."""
    import flipengine3d_amd as fa
    code = ('#include <stdio.h>\n#include "flipmi.h"\nint main(void){printf("%zu %zu %zu %zu\\n",'
            'sizeof(flipmi_scene_desc),sizeof(flipmi_frame_stats),sizeof(flipmi_stage_times),'
            'sizeof(flipmi_pressure_report));printf("%zu\\n",sizeof(flipmi_pressure_args));'
            'printf("%zu %zu %zu %zu\\n",sizeof(flipmi_server_desc),sizeof(flipmi_request),'
            'sizeof(flipmi_request_report),sizeof(flipmi_server_stats));'
            'printf("%zu\\n",sizeof(flipmi_surface_desc));return 0;}\n')
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "sizes.c")
        exe = os.path.join(d, "sizes")
        open(src, "w").write(code)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes == [ctypes.sizeof(fa.SceneDesc), ctypes.sizeof(fa.FrameStats),
                     ctypes.sizeof(fa.StageTimes), ctypes.sizeof(fa.PressureReport),
                     ctypes.sizeof(fa.PressureArgs), ctypes.sizeof(fa.ServerDesc), ctypes.sizeof(fa.Request),
                     ctypes.sizeof(fa.RequestReport), ctypes.sizeof(fa.ServerStats), ctypes.sizeof(fa.SurfaceDesc)]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_gpu_means_an_error_not_a_fallback(lib_path):
    import flipengine3d_amd as fa
    with pytest.raises(fa.FlipError):
        fa.Context(0)


def test_missing_library_raises(tmp_path):
    import flipengine3d_amd as fa
    saved = fa._lib
    fa._lib = None
    try:
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            fa.load_library(str(tmp_path / "libflipmi.so"))
    finally:
        fa._lib = saved


def test_cpp_engine_classes_compile_as_cpp14(tmp_path):
    # include/flipmi_engine.hpp (the reference's class signatures over the
    # C ABI) builds as C++14, the reference engine's language level, and
    # calls only entry points the C header declares.
    import subprocess
    src = tmp_path / "use.cpp"
    src.write_text('#include "flipmi_engine.hpp"\n'
                   "int main() { flipmi::PressureSolver p; flipmi::VelocityAdvector v; flipmi::ParticleAdvector a;\n"
                   "  flipmi::FluidSimulation s(8, 8, 8, 0.125); return p.getSolverStatus().size(); }\n")
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    text = open(os.path.join(ROOT, "include", "flipmi_engine.hpp")).read()
    used = set(re.findall(r"\b(flipmi_[a-z_0-9]+)\s*\(", text))
    assert used and used <= set(declared()), used - set(declared())
