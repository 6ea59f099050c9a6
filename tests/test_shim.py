"""The drop-in shims (ssenss_amd/shim/) compile against the reference's own
headers, define the reference entry points they replace, and call only
functions that libssx.so exports.  Skipped where the reference sources or
an MPI header are not mounted (the shims include the reference's
structs.h, which includes mpi.h)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
SHIM = os.path.join(ROOT, "ssenss_amd", "shim")
REF_SRC = "/root/reference/src"
MPI_INC = "/opt/conda/include"
LIB = os.path.join(ROOT, "ssenss_amd", "libssx.so")

# file -> reference functions it defines (file:line they replace)
DEFINES = {
    "ssx_shim.c": ["solve_poisson_cg",          # solver.c:159-173
                   "solve_helmholtz_cg",        # solver.c:175-189
                   "ssx_shim_init"],            # after main.c:79
    "ssx_particles_shim.c": ["compute_reference_domain_velocity",  # geometry.c:188-198
                             "integrate_particle_equations"],      # particles.c:182-218
}


def _symbols(path, dynamic=False):
    out = subprocess.run(["nm"] + (["-D"] if dynamic else []) + [path],
                         check=True, capture_output=True, text=True).stdout
    defined, undefined = set(), set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[-2] == "U":
            undefined.add(parts[-1])
        elif len(parts) == 3 and parts[1] in "TBDR":
            defined.add(parts[2])
    return defined, undefined


@pytest.mark.skipif(not (os.path.isdir(REF_SRC)
                         and os.path.isfile(os.path.join(MPI_INC, "mpi.h"))
                         and shutil.which("gcc")),
                    reason="reference sources or mpi.h not mounted")
@pytest.mark.parametrize("name", sorted(DEFINES))
def test_shim_compiles_against_reference_headers(name, tmp_path):
    obj = tmp_path / (name + ".o")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-c",
                    "-I", REF_SRC, "-I", os.path.join(ROOT, "include"),
                    "-I", MPI_INC, os.path.join(SHIM, name), "-o", str(obj)],
                   check=True)
    defined, undefined = _symbols(str(obj))
    for fn in DEFINES[name]:
        assert fn in defined, fn
    called = {s for s in undefined if s.startswith("ssx_")}
    assert called, "the shim calls the engine"
    if os.path.isfile(LIB):
        exported, _ = _symbols(LIB, dynamic=True)
        assert called <= exported, called - exported
