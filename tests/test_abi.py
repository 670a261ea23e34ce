"""The C-ABI library loads on a machine without a GPU and exports every entry
point that include/mdgpu/*.h declares (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "mdgpu")


def declared(header, pattern):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(pattern, text, flags=re.M)))


def test_c_abi_symbols_exported():
    from md_gpu_amd import engine
    lib = engine.lib()
    names = declared("md_api.h", r"^\s*(?:int|const char\s*\*)\s+(md_\w+)\s*\(")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python wrapper binds exactly these names
    assert set(engine.exported_symbols()) == set(names)


def test_compat_api_symbols_exported():
    from md_gpu_amd import engine
    names = declared("simulation.h", r"^\s*(?:simulation\s*\*|void|int|ffield\s*\*|config\s*\*)\s*(\w+)\s*\(")
    assert {"SimulationInit", "SimulationRun", "SimulationFinish", "SimulationRun_CUDA", "FFieldInput",
            "ConfigInput", "ConfigFinish", "FFieldFinish", "SimulationLastStatus"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\b_Z{len(n)}{n}\w*", out), n


def test_version_string_without_gpu():
    from md_gpu_amd import engine
    v = engine.version()
    assert "gfx950" in v


def test_athom_binary_built():
    path = os.path.join(ROOT, "md_gpu_amd", "bin", "athom")
    assert os.access(path, os.X_OK)
    r = subprocess.run([path], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


def test_engine_requires_a_device():
    """On a host without a GPU the engine refuses to start with a HIP error
    code instead of falling back to the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from md_gpu_amd import engine
    with pytest.raises(engine.MDError):
        engine.Engine(0)
