"""CPU checks of the drop-in boundary: the library exports every symbol the
public header declares (no compute calls), and the header compiles as C and
as C++."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hommexx_amd.h")
LIB = os.path.join(ROOT, "hommexx_amd", "libhommexx_amd.so")

REFERENCE_SYMBOLS = [
    "initialize_hommexx_session", "finalize_hommexx_session", "reset_cxx_comm", "init_connectivity",
    "add_connection", "finalize_connectivity", "init_derivative_c", "init_simulation_params_c",
    "init_elements_2d_c", "init_elements_states_c", "init_diagnostics_c", "init_hvcoord_c",
    "init_boundary_exchanges_c", "init_time_level_c", "prim_run_subcycle_c", "cxx_push_results_to_f90",
    "cxx_push_forcing_to_f90", "f90_push_forcing_to_cxx",
]


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = "\n".join(ln for ln in text.splitlines() if not ln.lstrip().startswith("#"))
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+[\s*]+(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_reference_abi():
    decl = declared_functions()
    for s in REFERENCE_SYMBOLS:
        assert s in decl, s


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libhommexx_amd.so not built")
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB]).decode()
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [s for s in declared_functions() if s not in exported]
    assert not missing, missing
    ctypes.CDLL(LIB)  # loads without a GPU (no compute call is made)


@pytest.mark.parametrize("lang,compiler", [("c", "gcc"), ("c++", "g++")])
def test_header_compiles(tmp_path, lang, compiler):
    src = tmp_path / ("t." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "hommexx_amd.h"\nint main(void) { return 0; }\n')
    subprocess.check_call([compiler, "-fsyntax-only", "-I", os.path.dirname(HEADER), str(src)])


def test_configure_bounds_nlev():
    """hommexx_hip_configure accepts 3 <= nlev <= 256 (the remap keeps each
    element's target table in LDS) and touches no GPU (separate process, so
    the library's session state stays private)."""
    if not os.path.exists(LIB):
        pytest.skip("libhommexx_amd.so not built")
    code = ("import ctypes, sys; L = ctypes.CDLL(sys.argv[1]); "
            "print([L.hommexx_hip_configure(n, 4) for n in (2, 3, 72, 256, 257, 4096)])")
    out = subprocess.check_output([os.sys.executable, "-c", code, LIB]).decode().strip()
    assert out == "[2, 0, 0, 0, 2, 2]", out


def _dims(lib):
    code = ("import ctypes, sys; L = ctypes.CDLL(sys.argv[1]); n = ctypes.c_int(); q = ctypes.c_int(); "
            "L.hommexx_hip_dimensions(ctypes.byref(n), ctypes.byref(q)); print(n.value, q.value)")
    out = subprocess.check_output([sys.executable, "-c", code, lib]).decode().split()
    return int(out[0]), int(out[1])


def test_build_dimensions_default_like_the_reference():
    """Without hommexx_hip_configure the library uses the build's PREQX_PLEV
    and QSIZE_D, whose defaults are the reference's (72 and 4,
    src/preqx/CMakeLists.txt:244-261)."""
    if not os.path.exists(LIB):
        pytest.skip("libhommexx_amd.so not built")
    assert _dims(LIB) == (72, 4)


def test_build_dimensions_follow_make_variables(tmp_path):
    """make PREQX_PLEV=26 QSIZE_D=6 (as cmake -DPREQX_PLEV=26 -DQSIZE_D=6
    builds the reference) gives a library whose defaults are 26 and 6."""
    lib = str(tmp_path / "libhx_L26_Q6.so")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "hommexx_amd", "csrc"), "-j8",
                           "BUILD=%s" % tmp_path, "LIB=%s" % lib, "PREQX_PLEV=26", "QSIZE_D=6"])
    assert _dims(lib) == (26, 6)
