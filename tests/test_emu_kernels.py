"""The kernels' control logic on the CPU: the library's own HIP sources compiled for the host
against a wave-emulation header (tests/emu), checked against the golden fixtures without a GPU.
Covers the split completer (items, device reduce against a host scan, budget-abort recursion,
capped items split deeper) and the stream kernel under forced donation. The product library
is never involved: the checks run in a subprocess that loads only the emulation build."""
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / 'emu'))
import build as emu_build  # noqa: E402


@pytest.fixture(scope='module')
def emu_lib():
    if emu_build.compiler() is None:
        pytest.skip('no host clang++ for the emulation build')
    return emu_build.build()


def test_kernel_logic_on_cpu_emulation_matches_golden(emu_lib):
    r = subprocess.run([sys.executable, '-u', str(HERE / 'emu' / 'run_checks.py')], capture_output=True, text=True,
                       timeout=900)
    out = r.stdout
    assert r.returncode == 0, out[-2000:] + r.stderr[-2000:]
    assert 'BAD' not in out, out
    assert out.strip().splitlines()[-1] == 'bad 0', out[-2000:]
    assert out.count('ok ') >= 10
