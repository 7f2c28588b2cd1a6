"""The KAPSEL-side C++ binding (kapsel3_amd/binding) against the
reference's own headers: it compiles with KAPSEL's variable.h (Particle,
CTime, exit_job), its seam function has the type of the `Time_evolution`
pointer as sp_3d_ns.cxx:10 declares it and its FFT shims the type of
A2a_k / A_k2a (fft_wrapper.h:641-681), it links against libkapsel_hip, and
every global its parameter translation reads is declared in input.h. CPU
only; skipped where the reference tree is absent."""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
BIND = os.path.join(ROOT, "kapsel3_amd", "binding")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="the reference tree is not here")


def ref_line(path, pattern):
    for line in open(os.path.join(REF, path)):
        if re.search(pattern, line):
            return line.strip()
    raise AssertionError(f"{pattern} not in {path}")


def test_binding_compiles_against_variable_h_and_matches_the_seam(tmp_path):
    seam = ref_line("sp_3d_ns.cxx", r"^void \(\*Time_evolution\)\(")          # sp_3d_ns.cxx:10
    fft = ref_line("fft_wrapper.h", r"inline void A2a_k\(double \*")          # fft_wrapper.h:641
    assert "double **zeta, double uk_dc[DIM], double **f, Particle *p, CTime &jikan" in seam
    src = tmp_path / "seam.cxx"
    src.write_text(f'''#include "variable.h"
#include "kapsel_binding.h"
{seam}
void (*fft_fwd)(double *a) = A2a_k_hip;
void (*fft_inv)(double *a) = A_k2a_hip;
int main() {{ Time_evolution = Time_evolution_hydro_hip; return fft_fwd == fft_inv; }}
''')
    inc = ["-I", REF, "-I", os.path.join(ROOT, "include"), "-I", BIND]
    for f in (str(src), os.path.join(BIND, "kapsel_binding.cxx")):
        p = subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", "-fsyntax-only", *inc, f],
                           capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
    assert "A2a_k" in fft


def test_binding_links_against_the_library(tmp_path):
    lib = os.path.join(ROOT, "kapsel3_amd", "libkapsel_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libkapsel_hip.so is not built")
    src = tmp_path / "main.cxx"
    src.write_text('''#include "variable.h"
#include "kapsel_binding.h"
int main(int argc, char **) { if (argc > 5) { double a[8]; A2a_k_hip(a); kh_binding_finish(); } return 0; }
''')
    inc = ["-I", REF, "-I", os.path.join(ROOT, "include"), "-I", BIND]
    exe = str(tmp_path / "a.out")
    p = subprocess.run(["g++", "-std=c++11", *inc, str(src), os.path.join(BIND, "kapsel_binding.cxx"),
                        "-L", os.path.dirname(lib), "-lkapsel_hip", "-Wl,-rpath-link,/opt/rocm/lib",
                        "-o", exe], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_parameter_translation_reads_declared_globals():
    """kapsel_binding_globals.cxx needs input.h, which needs OCTA's
    udfmanager.h (absent here): every global it reads must be declared
    extern in the reference's input.h."""
    text = open(os.path.join(BIND, "kapsel_binding_globals.cxx")).read()
    body = text[text.index("kh_params kh_params_from_globals"):text.index("/* In main")]
    used = {"Component_Number", "MASS_RATIOS", "janus_propulsion", "janus_axis", "janus_force", "janus_torque",
            "janus_slip_vel", "janus_slip_mode", "janus_rotlet_C1", "janus_rotlet_dipole_C2", "NX", "NY", "NZ",
            "DX", "RHO", "ETA", "kBT", "alpha_v", "alpha_o", "DT", "A", "A_XI", "G", "G_direction", "ROTATION",
            "FIX_CELLxyz", "LJ_truncate", "LJ_powers", "EPSILON", "MAX_SLIP_ITER", "MAX_SLIP_TOL"}
    for name in used:
        assert re.search(rf"\b{name}\b", body), name
    decl = open(os.path.join(REF, "input.h")).read()
    missing = [n for n in sorted(used) if not re.search(rf"^extern\s[^;]*\b{n}\b[^;]*;", decl, re.M)]
    assert not missing, missing
