"""Register and scratch budget of the default constitutive kernels, from
hipcc's resource-usage report (no GPU needed).

Private (scratch) memory in these kernels is HBM traffic: a zero-initialised
coefficient aggregate that the compiler left in scratch once added about 23
written bytes per point to the 424-byte stream and cost 12 % of the S+DS
throughput.  The default forms (tuning variant 11, with and without DS) must
keep everything in registers."""
import os
import re
import shutil
import subprocess

import pytest

from slope_stability_octave_amd import build

HIPCC = build.HIPCC if os.path.exists(build.HIPCC) else shutil.which("hipcc")


def resource_usage(src, tmp_path):
    """{kernel symbol: {counter: value}} from hipcc's resource-usage remarks."""
    flags = [f for f in build.FLAGS if f not in ("-shared",)]
    cmd = [HIPCC] + flags + ["-c", src, "-o", str(tmp_path / "k.o"), "-I", build.CSRC,
                             "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    usage, name = {}, None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            usage[name] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[bytes/lane\])?: (\d+)", line)
        if m and name:
            usage[name][m.group(1).strip()] = int(m.group(2))
    return usage


@pytest.mark.skipif(not HIPCC, reason="hipcc not available")
def test_default_constitutive_kernels_use_no_scratch(tmp_path):
    usage = resource_usage(os.path.join(build.CSRC, "constitutive.hip"), tmp_path)
    # k_mc3d_x36d<WITH_DS, NT=true, H=16, MINW=1, SDIRECT=false, MI> (both WITH_DS forms, per-point and
    # per-material fields) and the 2D kernels
    default = [n for n in usage if re.search(r"k_mc3d_x36dILb[01]ELb1ELi16ELi1ELb0ELb[01]E", n)]
    two_d = [n for n in usage if "k_mc2d" in n]
    assert len(default) == 4 and two_d, sorted(usage)
    for n in default + two_d:
        assert usage[n].get("ScratchSize") == 0, (n, usage[n])
        assert usage[n].get("VGPRs Spill", 0) == 0, (n, usage[n])


# alternative constitutive forms kept for the autotuner that hold more state
# than registers allow: the software-pipelined grid and the 5-waves-per-SIMD
# launch bounds (MINW = 5)
ALLOWED_SCRATCH = ("k_mc3d_pipe", "ELi5EEEvl", "ELi16ELi5E")


@pytest.mark.skipif(not HIPCC, reason="hipcc not available")
@pytest.mark.parametrize("src", ["assembly.hip", "krylov.hip", "mlp.hip", "constitutive.hip"])
def test_no_scratch_outside_the_alternative_forms(src, tmp_path):
    usage = resource_usage(os.path.join(build.CSRC, src), tmp_path)
    assert usage
    bad = {n: u.get("ScratchSize") for n, u in usage.items()
           if u.get("ScratchSize", 0) and not any(a in n for a in ALLOWED_SCRATCH)}
    assert not bad, bad
