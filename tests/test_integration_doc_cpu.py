"""The integration guide's C++ examples compile against include/fwi2d.h and
include/fwi2d_update.h (g++ -fsyntax-only), so the documented calls keep the
headers' signatures, and its ctypes stub matches the Python binding."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Declarations of the names the examples take from the surrounding
# reference classes (Modeling, Migration, Inversion members).
PRELUDE = r"""
#include "fwi2d.h"
#include "fwi2d_update.h"
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>
struct Geometry { int nsrc; };
void example(int nz, int nx, float dh, int nt, float dt, int nb, float abc_factor, float fmax,
             Geometry *geometry, int nsrc, const float *xsrc, const float *zsrc, int nrec,
             const float *xrec, const float *zrec, const float *Vp_nx_nz, const float *data,
             float *image_nx_nz, uint32_t iteration, float depth_mask, float *g_nx_nz,
             const double *alpha, const double *f, float *m_nx_nz, double m_min, double dm)
{
"""


def test_integration_cpp_examples_compile(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    assert len(blocks) >= 3
    body = "\n".join(b.replace('#include "fwi2d.h"', "").replace('#include "fwi2d_update.h"', "")
                     for b in blocks)
    src = tmp_path / "example.cpp"
    src.write_text(PRELUDE + body + "\n}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-unused-variable",
                        f"-I{os.path.join(ROOT, 'include')}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_integration_python_stub_matches_the_binding():
    """The guide's ctypes stub loads the library (no GPU call) and its
    fwi_config layout equals the binding's."""
    import ctypes as C
    from fwi2d_amd import abi
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```python\n(.*?)```", text, re.S)
    assert blocks
    ns = {}
    exec(blocks[0].replace("<repo>", ROOT), ns)
    stub, ref = ns["Config"], abi.Config
    assert C.sizeof(stub) == C.sizeof(ref)
    assert [(n, t) for n, t in stub._fields_] == [(n, t) for n, t in ref._fields_]
    for name, _ in ref._fields_:
        assert getattr(stub, name).offset == getattr(ref, name).offset, name
    assert callable(ns["check"]) and ns["lib"].fwi_create is not None
