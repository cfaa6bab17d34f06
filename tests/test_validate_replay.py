"""Memory-safety replay of the benchmark's whole call sequence on the CPU (no GPU).

The validation build of the library (-DRR_VALIDATE, rerdmft_amd/csrc/rr_validate.hpp) is linked
against a host-memory stand-in for the HIP runtime and the solver libraries
(tests/validate/hip_mock.cpp); instead of launching, every kernel launch site checks that each
pointer argument lies in a live allocation, every GEMM checks the highest element its operands and
outputs can touch against the allocation, and every caller array is checked at the extent the
entry point's contract gives it. tests/validate/replay.cpp replays bench.py's call sequence at its
full sizes (large real workload, uneven batch, sigma densities, X2C / RKB, then the per-row leg).

Regression: the per-row leg used to pass an n x n coefficient matrix to rr_ri_mo_fock_nonrel, whose
contract is n x 2n (spin orbitals, RiNonRelFullOptFock.cpp:53-98); the library then read 2 MB past
the caller's allocation, which faults the GPU whenever that memory is unmapped.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def replay_exe():
    from rerdmft_amd import build
    return build.build_validation()


def run(exe, legs, **env):
    e = dict(os.environ)
    e.pop("RRV_CONTROL", None)
    e.update({k: str(v) for k, v in env.items()})
    p = subprocess.run([exe, legs], capture_output=True, text=True, env=e, timeout=600)
    return p.returncode, p.stdout, p.stderr


def summary(out):
    last = [ln for ln in out.splitlines() if ln.startswith("replay:")][-1]
    parts = last.replace("replay:", "").split(",")
    return {p.split()[-1]: int(p.split()[0]) for p in parts}


def test_bench_call_sequence_is_memory_safe(replay_exe):
    rc, out, err = run(replay_exe, "main,batch,sigma,rel,rows,noalloc")
    s = summary(out)
    assert s["violations"] == 0, err[-3000:]
    assert s["calls"] == 0 and rc == 0, out[-3000:]
    assert s["checks"] > 5000  # every launch of the sequence went through the hooks


def test_short_spin_orbital_coefficients_are_reported(replay_exe):
    # the per-row leg's former call: c_current n x n instead of n x 2n
    rc, out, err = run(replay_exe, "rows", REPLAY_C_CURRENT_COLS=500)
    s = summary(out)
    assert rc == 1 and s["violations"] == 4
    assert err.count("caller input array (declared extent)") == 4
    assert "exceeds the allocation" in err and "by 2000000 bytes" in err


def test_planted_caller_mistakes_are_reported(replay_exe):
    rc, out, err = run(replay_exe, "plant")
    assert rc == 1
    assert "points into a freed allocation" in err
    assert "caller input array (declared extent)" in err


def test_python_binding_rejects_short_arrays():
    import rerdmft_amd as rr
    ctx = rr.Context.__new__(rr.Context)  # no device needed: the check runs before the C call
    ctx._h = None
    n = 50
    with pytest.raises(ValueError, match="c_current"):
        ctx.ri_mo_fock_nonrel(0, np.eye(n), np.eye(n), np.ones(2 * n), np.eye(2 * n), np.eye(2 * n))
    with pytest.raises(ValueError, match="v_total"):
        ctx.rkb_fock_ri(0, 1, 2, 3, np.eye(8), np.eye(8), np.eye(8), np.eye(8), np.eye(8))
    with pytest.raises(ValueError, match="out"):
        ctx.nonrel_fock_ri(0, np.eye(n), np.eye(n), np.eye(n), np.eye(n), out=np.empty((n, n - 1)))
