"""The reference-named C++ host layer (include/ctdm_ref.hpp) compiles, links against libctdm.so and, on
a GPU, reproduces the one-call batched pipeline when driven like the reference's per-frame loop."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpp" / "ref_api_demo.cpp"


def _build(out: Path) -> Path:
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", str(ROOT / "include"), "-I", "/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", str(SRC), "-o", str(out), "-L", str(ROOT / "ctdm_amd"), "-lctdm",
           "-L", "/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{ROOT / 'ctdm_amd'}", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_cpp_layer_compiles_and_links(tmp_path):
    assert _build(tmp_path / "demo").exists()


@pytest.mark.gpu
def test_cpp_layer_matches_pipeline(tmp_path):
    exe = _build(tmp_path / "demo")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK")
