"""A C99 host program (tests/c_abi/host_demo.c) compiled against
include/openqcd_hip.h with gcc -std=c99 -Wall -Wextra -Werror and linked
against libopenqcd_hip.so: the header is valid C, every function it
declares is exported by the library (the program takes the address of each
one), and, on the GPU, a reference-style solve through the C calls reaches
its true residual."""
import os
import subprocess

import pytest

import pyoqh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_abi", "host_demo.c")
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "openqcd-2.4.2_amd")


def _build(tmp):
    names = pyoqh.header_functions()
    with open(os.path.join(tmp, "host_demo_symbols.h"), "w") as f:
        f.write("static void *const oqh_symbols[] = {\n")
        for n in names:
            f.write("   (void *)&%s,\n" % n)
        f.write("};\n")
    exe = os.path.join(tmp, "host_demo")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror",
           "-Wstrict-prototypes", "-I" + INC, "-I" + tmp, SRC, "-o", exe, "-L" + LIBDIR,
           "-Wl,-rpath," + LIBDIR, "-l:libopenqcd_hip.so", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe, len(names)


def test_header_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libopenqcd_hip.so")):
        pytest.skip("library not built")
    exe, n = _build(str(tmp_path))
    r = subprocess.run([exe, "symbols"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["symbols", str(n)]


@pytest.mark.gpu
def test_c_host_solve(tmp_path):
    exe, _ = _build(str(tmp_path))
    r = subprocess.run([exe, "solve"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    f = r.stdout.split()
    assert f[0] == "status" and int(f[3]) == 0 and float(f[7]) <= 1.2e-10
