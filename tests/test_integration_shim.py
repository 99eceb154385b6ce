"""The reference-side binding in INTEGRATION.md compiles against the
reference's own mesh.h, particle.h and constants.h (with the image's MPICH
headers): every Domain field, enum and reference signature it names exists
with the type it is used as. CPU only; skipped where the reference sources
are absent (they never travel with the build)."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
MPI_INC = Path("/opt/conda/include")


def _blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


@pytest.mark.skipif(not (REF / "mesh.h").exists() or shutil.which("g++") is None or not (MPI_INC / "mpi.h").exists(),
                    reason="reference headers, g++ or MPI headers absent")
def test_shim_compiles_against_the_reference_headers(tmp_path):
    b = _blocks()
    assert len(b) == 4
    shim, loop, hdf, jumps = b
    tu = (shim + "\n#include <cstdio>\n" + hdf +
          "\nvoid felis_loop(Domain &D, int it0, int n, int wake_step, double *energy, double *bfac,"
          " double *twiss) {\n" + loop + jumps + "}\n")
    src = tmp_path / "felis_shim.cpp"
    src.write_text(tu)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wno-unused-variable",
                        "-Wno-unused-parameter", f"-I{REF}", f"-I{ROOT / 'include'}", f"-I{MPI_INC}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
