"""The C++ host layer (include/mtimager_amd_gafw.hpp): GAFW-style operators
over the C-ABI, driven by tests/host/gafw_check (built by
__graft_entry__.build() or `make -C tests/host`)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "host", "gafw_check")


def _exe():
    if not os.path.exists(EXE):
        r = subprocess.run(["make", "-C", os.path.join(HERE, "host")], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    return EXE


def test_operators_validate_like_the_reference():
    """Factory names, property parsing, output shapes/types and the
    ValidationException split -- no GPU involved."""
    r = subprocess.run([_exe(), "validate"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout


@pytest.mark.gpu
def test_operators_image_like_the_oracle():
    """ConvolutionGridder (two accumulated chunks in the reference's WImager
    layout) and ImageFinalizer through the host layer vs the FP64 oracle."""
    r = subprocess.run([_exe(), "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
