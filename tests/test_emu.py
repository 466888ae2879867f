"""The HIP sources built for the CPU against the fibre-based HIP emulation in tests/emu (g++,
AddressSanitizer): the radix sort and scan unit driver, and two Newton steps of a two-cube scene
through the C ABI compared with the oracle (indexing errors surface here as sanitizer reports,
before any GPU run), then two steps with the learned step predictor (body features against the
oracle's, the step against the oracle given the same prediction), two steps with the
multi-tile persistent PCG, and the assembly's PSD projection on the matrix cores (the instruction
emulated with the operand layout the kernels assume; test_gpu_features checks it on the hardware)
against eigendecompositions."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))

import build_emu  # noqa: E402


@pytest.fixture(scope="module")
def emu_lib():
    return build_emu.build()


def test_sort_and_scan_driver():
    exe = build_emu.build_sort_driver()
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = out.stdout.strip().splitlines()
    assert lines and all(line.endswith("ok") for line in lines), out.stdout[-2000:]


def test_emulated_step_matches_oracle(emu_lib):
    env = dict(os.environ, LD_PRELOAD=build_emu.asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    out = subprocess.run([sys.executable, os.path.join(HERE, "emu", "emu_step.py")], capture_output=True, text=True,
                         env=env, timeout=900)
    assert out.returncode == 0, (out.stdout[-2000:], out.stderr[-4000:])
    res = [line for line in out.stdout.splitlines() if line.startswith("result")]
    assert res, out.stdout[-2000:]
    _, err, na_emu, na_orc = res[-1].split()
    assert float(err) < 1e-6
    assert int(na_emu) == int(na_orc)
    mul = [line for line in out.stdout.splitlines() if line.startswith("multi")]
    assert mul and float(mul[-1].split()[1]) < 1e-6, out.stdout[-2000:]
    psd = [line for line in out.stdout.splitlines() if line.startswith("psd")]
    assert psd, out.stdout[-2000:]
    _, e64, e32 = psd[-1].split()
    assert float(e64) < 1e-13 and float(e32) < 2e-6
    lrn = [line for line in out.stdout.splitlines() if line.startswith("learned")]
    assert lrn, out.stdout[-2000:]
    _, ferr, serr, pa_emu, pa_orc = lrn[-1].split()
    assert float(ferr) < 1e-5 and float(serr) < 1e-6
    assert float(pa_emu) == float(pa_orc)
