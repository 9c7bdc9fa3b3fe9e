"""No gfx950 kernel of the library spills to scratch memory (checked from the
compiler's kernel-resource-usage remarks; no GPU needed)."""
import os
import subprocess
import sys

from conftest import PKG


def test_no_kernel_uses_scratch():
    r = subprocess.run([sys.executable, os.path.join(PKG, "check_resources.py")], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "none uses scratch memory" in r.stdout
