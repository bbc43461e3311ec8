"""Host-side C++ (difx-svn2git-try3_amd/host/) checks that need no GPU."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(os.path.dirname(HERE), "difx-svn2git-try3_amd", "_build")


def test_delay_model_interpolator():
    exe = os.path.join(BUILD, "test_delay_model")
    assert os.path.exists(exe), "build first (__graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


def test_two_bit_transpose():
    """The register transpose of the 2-bit corner turn (csrc/fx_bits.h)."""
    exe = os.path.join(BUILD, "test_bits")
    assert os.path.exists(exe), "build first (__graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
