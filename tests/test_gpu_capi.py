"""The C-ABI driven from C: tests/capi/bsx_capi_check (built by
__graft_entry__.build()) creates, inserts into, finishes and destroys a
reconstruction and extracts central sections through include/bsx.h, and
compares each result with the FP64 oracle on the same seeded inputs."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "capi", "bsx_capi_check")


@pytest.mark.gpu
def test_c_program_against_oracle(gpu):
    assert os.path.exists(EXE), "tests/capi/bsx_capi_check is not built (run build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout


@pytest.mark.gpu
def test_dropin_core_against_oracle(gpu):
    """bsoft_amd/dropin/bsx_dropin.hpp as breconstruct drives
    particle_reconstruct and img_reconstruction_sum_weigh: 4 round-robin
    partial maps, every output map of nmaps 1, 2 and 3 against the oracle."""
    exe = os.path.join(HERE, "capi", "bsx_dropin_check")
    assert os.path.exists(exe), "tests/capi/bsx_dropin_check is not built (run build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
