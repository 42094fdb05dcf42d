"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/nmc.h declares; host-side validation paths that
need no device."""
import ctypes
import os
import re

from tests.conftest import ROOT


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "nmc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nmc_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from nanomagmc_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_python_binding_signatures_cover_header():
    from nanomagmc_amd import _lib
    L = _lib.lib()
    for s in header_symbols():
        assert getattr(L, s).argtypes is not None or s == "nmc_last_error"


def test_bond_parser_matches_oracle(orc, bloch3d):
    from nanomagmc_amd import parse_bonds
    b = parse_bonds(bloch3d, 3, 1.0, 0.5)
    d, J, D = orc.parse_bonds(bloch3d, 3, 1.0, 0.5)
    assert [x.d for x in b] == [tuple(r) for r in d.tolist()]
    assert [x.J for x in b] == J.tolist()
    assert [list(x.D) for x in b] == D.tolist()


def test_crop_to_box():
    import numpy as np
    from nanomagmc_amd import crop_to_box
    m = np.zeros((6, 6), dtype=np.uint8)
    m[2:4, 1:5] = 1
    n, lo, sub = crop_to_box(m.reshape(-1), 2, 6)
    assert n == [2, 4] and lo == [2, 1] and sub.sum() == 8


def test_device_index_math_on_host(tmp_path):
    """Fast division and the XCD block->tile permutation (compiled for the
    host from the device header)."""
    import subprocess
    exe = str(tmp_path / "index_check")
    src = os.path.join(ROOT, "tests", "index_check.cpp")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", src, "-o", exe], check=True,
                   capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "fdiv_bad=0 xcd_bad=0" in r.stdout


def test_header_is_plain_c_and_links(tmp_path):
    """include/nmc.h compiles as C99 with warnings as errors, a C program
    taking the address of every declared entry point links against
    libnmc.so, and it runs (calling only nmc_last_error, which needs no
    device)."""
    import subprocess
    from nanomagmc_amd import _lib
    syms = header_symbols()
    src = tmp_path / "use_nmc.c"
    src.write_text(
        '#include <stdio.h>\n#include "nmc.h"\n'
        "typedef void (*fn)(void);\n"
        "static fn table[] = {\n" + ",\n".join("    (fn)%s" % s for s in syms) + "\n};\n"
        "int main(void) {\n"
        "    const char* e = nmc_last_error();\n"
        "    printf(\"%d %s\\n\", (int)(sizeof(table) / sizeof(table[0])), e ? \"ok\" : \"null\");\n"
        "    return 0;\n}\n")
    exe = str(tmp_path / "use_nmc")
    libdir = os.path.dirname(_lib.LIB_PATH)
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-Wno-cast-function-type",
                        "-I", os.path.join(ROOT, "include"), str(src), "-o", exe, "-L", libdir, "-lnmc",
                        "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == [str(len(syms)), "ok"]
