"""The C++ adapter (rerdmft_amd/host/rerdmft_gpu.hpp) as a drop-in: a templated RHF SCF loop run with
the GPU ERI source and with the CPU oracle must reach the same energy (1e-10 Ha); the same loop with
the adapter's GpuDiis / scfCommutatorError must converge to the linearly mixed loop's energy in fewer
iterations."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "rerdmft_amd", "_build", "test_adapter")
ORACLE = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


@pytest.mark.gpu
def test_cpp_adapter_scf_matches_oracle():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_adapter.cpp was not built (python -m rerdmft_amd.build)")
    r = subprocess.run([EXE, ORACLE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[PASS]" in r.stdout
    assert "density_matrix dimension does not match nl" in r.stdout
    assert "E_diis" in r.stdout
