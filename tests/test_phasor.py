"""The device phasor (FP64 reduction, FP32 Taylor polynomials) checked on the host against
long-double sin / cos: maximum error below 3e-7 for phases up to 1e8 rad (tests/phasor_check.cpp,
compiled from the same header as the kernels)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "ddfacet_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_phasor_accuracy_on_host():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "phasor_check")
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-I", CSRC,
                        os.path.join(HERE, "phasor_check.cpp"), "-o", exe], check=True,
                       capture_output=True, timeout=300)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
