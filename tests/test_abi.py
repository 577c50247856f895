"""The C ABI library loads, exports every entry point include/ctdm.h declares, and the Python
binding's signature table covers exactly those (no compute calls: runs without a GPU)."""
import re
import subprocess
from pathlib import Path

import ctdm_amd

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "ctdm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(ctdm_[a-z0-9_]+)\s*\(", text))


def test_header_declarations_exported():
    lib = ctdm_amd.lib()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(ctdm_amd.LIBPATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (ctdm_[a-z0-9_]+)", out))
    assert names <= exported


def test_binding_table_matches_header():
    assert set(ctdm_amd.SIGNATURES) == declared()


def test_version_and_null_context():
    lib = ctdm_amd.lib()
    assert b"gfx950" in lib.ctdm_version()
    assert lib.ctdm_last_error(None) == b"null context"
    assert lib.ctdm_sync(None) == ctdm_amd.EINVAL


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "ctdm.h"\nint main(void){ ctdm_crop c; (void)c; return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(ROOT / "include"), "-c",
                    str(src), "-o", str(tmp_path / "t.o")], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", str(ROOT / "include"), "-x",
                    "c++", "-c", str(src), "-o", str(tmp_path / "t2.o")], check=True)
