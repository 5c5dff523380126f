"""The library's kernel and runtime sources, compiled for the CPU with
AddressSanitizer and UBSan (tests/emu) and run on tiny meshes against the
oracle: every device access is bounds-checked, LDS starts as garbage for
each block, and graph capture refuses memset/copy/event nodes.  The check
runs in a child process because the sanitizer runtime must be preloaded."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))


def _asan():
    if not shutil.which("g++"):
        return None
    p = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(_asan() is None, reason="g++ with AddressSanitizer not available")
def test_emulated_kernels_match_oracle():
    import build as emu_build

    emu_build.build()
    env = dict(os.environ, LD_PRELOAD=_asan(), ASAN_OPTIONS="detect_leaks=0", EMU_POISON="1")
    p = subprocess.run([sys.executable, os.path.join(HERE, "emu", "drive.py")], env=env,
                       capture_output=True, text=True, timeout=1200)
    print(p.stdout[-3000:])
    print(p.stderr[-3000:])
    assert p.returncode == 0 and "ALL PASS" in p.stdout
