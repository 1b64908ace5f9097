"""The ISO_C_BINDING shim compiles against the reference's `module global` and its
sofia_params_t matches the C struct (runs where amdflang and oracle/_ref are present)."""
import ctypes
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "molecular_simulation_code_amd", "fortran", "sofia_hip_iface.f90")
REFMOD = os.path.join(ROOT, "oracle", "_ref", "obj")


def _fc():
    for c in (shutil.which("amdflang"), "/opt/rocm/bin/amdflang"):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    fc = _fc()
    if fc is None or not os.path.exists(os.path.join(REFMOD, "global.mod")):
        pytest.skip("amdflang or the reference's global.mod (oracle/build_ref.sh) not available")
    d = tmp_path_factory.mktemp("fshim")
    return fc, str(d)


def test_shim_compiles_against_reference_global(workdir):
    fc, d = workdir
    r = subprocess.run([fc, "-c", "-w", "-I", REFMOD, "-module-dir", d, SHIM, "-o",
                        os.path.join(d, "shim.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_params_type_matches_c_struct(workdir):
    fc, d = workdir
    from molecular_simulation_code_amd._lib import SofiaParams
    subprocess.run([fc, "-c", "-w", "-I", REFMOD, "-module-dir", d, SHIM, "-o",
                    os.path.join(d, "shim.o")], check=True, capture_output=True)
    prog = os.path.join(d, "sz.f90")
    with open(prog, "w") as f:
        f.write("program sz\n use iso_c_binding\n use sofia_hip_iface\n type(sofia_params_t) :: p\n"
                " print *, c_sizeof(p)\nend program\n")
    exe = os.path.join(d, "sz")
    # linking also proves every C entry point the shim binds exists in libsofia_hip.so
    libdir = os.path.join(ROOT, "molecular_simulation_code_amd")
    r = subprocess.run([fc, "-w", "-I", d, prog, os.path.join(d, "shim.o"),
                        os.path.join(REFMOD, "global.o"), "-L", libdir, "-lsofia_hip",
                        f"-Wl,-rpath,{libdir}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    assert int(out.split()[0]) == ctypes.sizeof(SofiaParams)
