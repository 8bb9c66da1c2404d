"""The MEX drop-in sources (slope_stability_octave_amd/mex/*.c) compile
against include/ssr_hip.h and a declaration-only MEX header (Octave is not
installed here), each defines mexFunction, and together they cover the seven
MEX files of the reference's hot path (CPU only)."""
import glob
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MEX = os.path.join(ROOT, "slope_stability_octave_amd", "mex")
REFERENCE_MEX = {"constitutive_problem_3D_SDS_mex", "constitutive_problem_3D_S_mex", "constitutive_problem_2D_SDS_mex",
                 "constitutive_problem_2D_S_mex", "assemble_K_tangent_vals", "assemble_K_tangent_vals_2D",
                 "hypre_boomeramg_mex"}


def test_every_reference_mex_has_a_shim():
    names = {os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(MEX, "*.c"))}
    assert names == REFERENCE_MEX


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc missing")
@pytest.mark.parametrize("name", sorted(REFERENCE_MEX))
def test_shim_compiles_against_the_c_abi(name):
    src = os.path.join(MEX, name + ".c")
    text = open(src).read()
    assert re.search(r"void\s+mexFunction\s*\(\s*int\s+nlhs", text)
    assert re.search(r"\bssr_\w+\s*\(", text)
    out = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
                          "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tests", "mexstub"), src],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
