"""The C-ABI boundary, checked without a GPU: the library loads, exports
every function include/*.h declares, and the public structs have the
reference's x86-64 layout (src/bspline_base.h:49-96,
src/multi_bspline_structs.h:29-180).  The expected offsets are worked out
from the reference's definitions (4-byte enums and ints, 8-byte pointers,
intptr_t and doubles, natural alignment)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in sorted(os.listdir(INC)):
        if not f.endswith(".h"):
            continue
        src = open(os.path.join(INC, f)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = re.sub(r"#[^\n]*", "", src)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(", src):
            name = m.group(1)
            # a declaration: name( ... ) followed by ';' at top level
            rest = src[m.end():]
            depth, i = 1, 0
            while depth and i < len(rest):
                depth += {"(": 1, ")": -1}.get(rest[i], 0)
                i += 1
            if rest[i:].lstrip().startswith(";") and name not in ("sizeof",):
                names.add(name)
    return names


def test_library_exports_every_declared_function():
    import einspline_amd as ea  # noqa: F401  (raises if the .so is missing)
    lib = os.path.join(ROOT, "einspline_amd", "libeinspline.so")
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = declared_functions()
    assert len(declared) > 50
    missing = sorted(declared - exported)
    assert not missing, missing


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "multi_bspline.h"
#define P(T, F) printf(#T "." #F " %zu\n", offsetof(T, F));
#define S(T) printf(#T ".sizeof %zu\n", sizeof(T));
int main(void) {
  S(Ugrid) P(Ugrid, num) P(Ugrid, delta) P(Ugrid, delta_inv)
  S(BCtype_s) S(BCtype_d) S(BCtype_c) S(BCtype_z) S(Bspline) P(Bspline, coefs)
  S(multi_UBspline_3d_s) P(multi_UBspline_3d_s, x_grid) P(multi_UBspline_3d_s, xBC)
  P(multi_UBspline_3d_s, num_splines)
  S(multi_UBspline_3d_d) P(multi_UBspline_3d_d, zBC) P(multi_UBspline_3d_d, num_splines)
  S(multi_UBspline_3d_c) P(multi_UBspline_3d_c, num_splines) P(multi_UBspline_3d_c, lapl3)
  S(multi_UBspline_3d_z) P(multi_UBspline_3d_z, zBC) P(multi_UBspline_3d_z, lapl3)
  S(multi_UBspline_1d_s) P(multi_UBspline_1d_s, xBC) P(multi_UBspline_1d_s, num_splines)
  S(multi_UBspline_1d_d) P(multi_UBspline_1d_d, num_splines)
  S(multi_UBspline_1d_c) P(multi_UBspline_1d_c, num_splines)
  S(multi_UBspline_1d_z) P(multi_UBspline_1d_z, num_splines)
  S(multi_UBspline_2d_s) P(multi_UBspline_2d_s, y_grid) P(multi_UBspline_2d_s, yBC)
  P(multi_UBspline_2d_s, num_splines)
  S(multi_UBspline_2d_d) P(multi_UBspline_2d_d, num_splines)
  S(multi_UBspline_2d_c) P(multi_UBspline_2d_c, num_splines) P(multi_UBspline_2d_c, lapl2)
  S(multi_UBspline_2d_z) P(multi_UBspline_2d_z, yBC) P(multi_UBspline_2d_z, lapl2)
  return 0;
}
"""

# x86-64 layout of the reference's definitions
EXPECTED = {
    "Ugrid.sizeof": 40, "Ugrid.num": 16, "Ugrid.delta": 24, "Ugrid.delta_inv": 32,
    "BCtype_s.sizeof": 16, "BCtype_d.sizeof": 24, "BCtype_c.sizeof": 24, "BCtype_z.sizeof": 40,
    "Bspline.sizeof": 16, "Bspline.coefs": 8,
    "multi_UBspline_3d_s.sizeof": 216, "multi_UBspline_3d_s.x_grid": 40,
    "multi_UBspline_3d_s.xBC": 160, "multi_UBspline_3d_s.num_splines": 208,
    "multi_UBspline_3d_d.sizeof": 240, "multi_UBspline_3d_d.zBC": 208,
    "multi_UBspline_3d_d.num_splines": 232,
    "multi_UBspline_3d_c.sizeof": 248, "multi_UBspline_3d_c.num_splines": 232,
    "multi_UBspline_3d_c.lapl3": 240,
    "multi_UBspline_3d_z.sizeof": 296, "multi_UBspline_3d_z.zBC": 240,
    "multi_UBspline_3d_z.lapl3": 288,
    # 1-D: {codes 0,4; coefs 8; x_stride 16; x_grid 24..64; xBC 64; num_splines}
    "multi_UBspline_1d_s.sizeof": 88, "multi_UBspline_1d_s.xBC": 64,
    "multi_UBspline_1d_s.num_splines": 80,
    "multi_UBspline_1d_d.sizeof": 96, "multi_UBspline_1d_d.num_splines": 88,
    "multi_UBspline_1d_c.sizeof": 96, "multi_UBspline_1d_c.num_splines": 88,
    "multi_UBspline_1d_z.sizeof": 112, "multi_UBspline_1d_z.num_splines": 104,
    # 2-D: {codes; coefs 8; strides 16, 24; grids 32, 72; BCs 112..; num_splines[; lapl2]}
    "multi_UBspline_2d_s.sizeof": 152, "multi_UBspline_2d_s.y_grid": 72,
    "multi_UBspline_2d_s.yBC": 128, "multi_UBspline_2d_s.num_splines": 144,
    "multi_UBspline_2d_d.sizeof": 168, "multi_UBspline_2d_d.num_splines": 160,
    "multi_UBspline_2d_c.sizeof": 176, "multi_UBspline_2d_c.num_splines": 160,
    "multi_UBspline_2d_c.lapl2": 168,
    "multi_UBspline_2d_z.sizeof": 208, "multi_UBspline_2d_z.yBC": 152,
    "multi_UBspline_2d_z.lapl2": 200,
}


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_struct_layout_matches_reference(lang, tmp_path):
    src = tmp_path / ("probe.c" if lang == "c" else "probe.cpp")
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    cc = "gcc" if lang == "c" else "g++"
    subprocess.run([cc, "-I", INC, str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True,
                                                         text=True, check=True).stdout.splitlines())
    got = {k: int(v) for k, v in got.items()}
    assert got == EXPECTED


def test_headers_compile_standalone(tmp_path):
    for h in sorted(os.listdir(INC)):
        for cc, ext in (("gcc", "c"), ("g++", "cpp")):
            f = tmp_path / f"t.{ext}"
            f.write_text(f'#include "{h}"\nint main(void) {{ return 0; }}\n')
            subprocess.run([cc, "-Wall", "-Werror", "-I", INC, "-fsyntax-only", str(f)], check=True)


def test_oracle_is_not_linked_into_the_product():
    lib = os.path.join(ROOT, "einspline_amd", "libeinspline.so")
    out = subprocess.run(["nm", "-D", lib], capture_output=True, text=True).stdout
    assert "oracle_" not in out
    deps = subprocess.run(["ldd", lib], capture_output=True, text=True).stdout
    assert "liboracle" not in deps


def test_cxx_callers_of_reference_headers_link(tmp_path):
    """A C++ object compiled against the reference's declarations (no
    extern "C", src/multi_bspline_eval_s.h:86-109) links against the
    library: the C++-mangled eval symbols exist."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include <complex>
#include "multi_bspline_structs.h"
void eval_multi_UBspline_3d_s_vgh(multi_UBspline_3d_s *, double, double, double, float *, float *, float *);
void eval_multi_UBspline_3d_z_vgl(multi_UBspline_3d_z *, double, double, double, std::complex<double> *,
                                  std::complex<double> *, std::complex<double> *);
int main(int argc, char **) {
  if (argc > 5) {   /* never called: linking is the test */
    eval_multi_UBspline_3d_s_vgh(0, 0, 0, 0, 0, 0, 0);
    eval_multi_UBspline_3d_z_vgl(0, 0, 0, 0, 0, 0, 0);
  }
  return 0;
}
''')
    libdir = os.path.join(ROOT, "einspline_amd")
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-I", INC, str(src), "-L", libdir, "-leinspline", "-Wl,-rpath," + libdir,
                    "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)], check=True)


def test_no_silent_cpu_fallback_without_a_gpu():
    """Without a visible GPU the product fails loudly: device tables cannot be
    created and the batched entry points return error codes (there is no CPU
    evaluation path in the library)."""
    import numpy as np

    import einspline_amd as ea
    if ea.device_count() > 0:
        pytest.skip("a GPU is visible; this checks the no-GPU behaviour")
    h = ea.HostSpline("s", (5, 5, 5), (0, 0, 0), (1, 1, 1), 3)
    h.set_block(0, np.random.default_rng(0).uniform(-1, 1, (3, 5, 5, 5)))
    with pytest.raises(ea.EinsplineError) as e:
        ea.DeviceTable(h)
    assert "hip" in str(e.value).lower()
    assert ea.lib.einspline_eval_vgh_batch(None, None, 1, None, None, None, 0, None) == ea.E_INVAL
    assert ea.lib.einspline_hip_malloc(16) is None


def test_autotune_entry_points_refuse_bad_arguments():
    """Both tuners check their arguments before any device work: no table,
    a mode outside v..vgh, no positions or no points are EINSPLINE_EINVAL."""
    import ctypes as C

    import numpy as np

    import einspline_amd as ea
    pts = np.zeros((4, 3))
    pp = pts.ctypes.data_as(C.POINTER(C.c_double))
    assert ea.lib.einspline_autotune(None, 3, 64, None, None, None, None) == ea.E_INVAL
    assert ea.lib.einspline_autotune_points(None, 3, pp, 4, None, None, None, None) == ea.E_INVAL
    fake = ea.multi_UBspline_3d_hip()
    assert ea.lib.einspline_autotune_points(C.byref(fake), 3, None, 4, None, None, None, None) == ea.E_INVAL
    assert ea.lib.einspline_autotune_points(C.byref(fake), 3, pp, 0, None, None, None, None) == ea.E_INVAL
    assert ea.lib.einspline_autotune_points(C.byref(fake), 5, pp, 4, None, None, None, None) == ea.E_INVAL


def test_every_declared_function_has_a_ctypes_prototype():
    """ctypes passes a Python int to an unprototyped function as a C int
    masked to 32 bits, which truncates device addresses: every function of
    include/*.h must be declared before the tests call it, and a call of an
    undeclared one raises instead of running."""
    import einspline_amd as ea
    import einspline_amd.serve  # noqa: F401  (declares the server functions)
    missing = sorted(n for n in declared_functions() if getattr(ea.lib, n).argtypes is None)
    assert not missing, missing
    fresh = ea.lib._FuncPtr(("einspline_hip_device_count", ea.lib))
    with pytest.raises(TypeError):
        fresh()
