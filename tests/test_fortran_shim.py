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


def test_restart_wrappers_round_trip_through_module_global(workdir, tmp_path):
    """dump_global_velocity_id_mpi_hip writes restart_<step> from module global and
    read_dump_global_id_v_hip reads it back into global, records placed by global id; the
    file also reads through the library's own reader."""
    fc, d = workdir
    subprocess.run([fc, "-c", "-w", "-I", REFMOD, "-module-dir", d, SHIM, "-o",
                    os.path.join(d, "shim.o")], check=True, capture_output=True)
    prog = os.path.join(d, "rs.f90")
    with open(prog, "w") as f:
        f.write(f"""program rs
 use global
 use sofia_hip_dropin
 implicit none
 integer :: i
 np = 3
 allocate(position(np), v(np), global_id(np))
 do i = 1, np
   position(i)%x = 1.5 * i; position(i)%y = 2.25 * i; position(i)%z = 40.0 - i
   v(i)%x = 0.001 * i; v(i)%y = -0.002 * i; v(i)%z = 0.0005
 end do
 global_id = (/ 3, 1, 2 /)
 box = 51.72; temp = 300.0; timeclock = 12.5
 restart_files_path = '{tmp_path}'
 call dump_global_velocity_id_mpi_hip(40, 9)
 position(:)%x = 0.0; box = 1.0
 call read_dump_global_id_v_hip('{tmp_path}/restart_40')
 print *, box, ibox, timeclock
 do i = 1, np
   print *, position(i)%x, position(i)%y, position(i)%z, v(i)%x
 end do
end program
""")
    exe = os.path.join(d, "rs")
    libdir = os.path.join(ROOT, "molecular_simulation_code_amd")
    r = subprocess.run([fc, "-w", "-I", d, "-I", REFMOD, prog, os.path.join(d, "shim.o"),
                        os.path.join(REFMOD, "global.o"), "-L", libdir, "-lsofia_hip",
                        f"-Wl,-rpath,{libdir}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    box, ibox, tc = (float(x) for x in out[0].split())
    assert box == pytest.approx(51.72, rel=1e-6) and ibox == pytest.approx(1 / 51.72, rel=1e-6)
    assert tc == pytest.approx(12.5)
    # atom k of the input had global id (3, 1, 2)[k]: slot g now holds that atom
    rows = [[float(x) for x in l.split()] for l in out[1:4]]
    for k, g in enumerate((3, 1, 2)):
        i = k + 1
        assert rows[g - 1] == pytest.approx([1.5 * i, 2.25 * i, 40.0 - i, 0.001 * i], abs=1e-6)
    from molecular_simulation_code_amd import io as sio
    b, t, num, tcl, x, v = sio.read_restart(tmp_path / "restart_40", 3)
    assert (num, b, t) == (9, pytest.approx(51.72), pytest.approx(300.0))
    assert x[2] == pytest.approx([1.5, 2.25, 39.0])
