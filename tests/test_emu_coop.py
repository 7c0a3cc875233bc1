"""The wavefront-cooperative qp kernels (plasticity_amd/csrc/cp_coop.h, cp_passes.h) on the CPU.

tests/emu builds the same kernel source for the host with one std::thread per lane (barriers for the
group's LDS syncs, ballots and shuffles) and AddressSanitizer on exact-size buffers, LDS slices
included.  So every access of the cooperative form is bounds-checked here, and a cross-lane LDS
dependency without a group sync shows up as a race.  Checks, each in a child process with the ASan
runtime preloaded:
  * the cooperative form equals the one-lane routine (cp_qp.h) bit for bit (FCC 12, HCP 18,
    BCC 48 systems);
  * per-qp results against the oracle at 1e-10 (batch form; with kinematic hardening; HCP);
  * the assembly and post-increment passes (persistent wavefronts on the work queue) against the
    oracle on a random displacement field and on the oracle's own states of fcc/simpleTension.
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def emu_env():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("no AddressSanitizer runtime (gcc) on this machine")
    r = subprocess.run(["make", "-s", "-C", EMU], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    return env


def _run(env, *args):
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_emu.py")] + [str(a) for a in args], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("case,strain,n", [("fcc", 1e-2, 48), ("hcp", 3e-3, 32), ("bcc48", 6e-4, 16)])
def test_coop_equals_one_lane_routine(emu_env, case, strain, n):
    out = _run(emu_env, "lane", case, strain, n, 3)
    assert out["na_mismatch"] == 0
    for k, v in out["coop_vs_lane"].items():
        assert v == 0.0, (k, v)


@pytest.mark.parametrize("case,strain,n,kin", [("fcc", 1e-2, 48, ""), ("fcc", 2e-3, 32, "kin"), ("hcp", 3e-3, 32, "")])
def test_coop_batch_matches_oracle(emu_env, case, strain, n, kin):
    out = _run(emu_env, "batch", case, strain, n, 11, *([kin] if kin else []))
    for flag in ("flag0", "flag1"):
        assert out[flag]["n_active_mismatch"] == 0, out
        for k, v in out[flag]["err"].items():
            assert v < 1e-10, (flag, k, v)
    assert out["flag1"]["plastic"] > n // 4


@pytest.mark.parametrize("kin", ["", "kin"])
def test_coop_passes_match_oracle(emu_env, kin):
    """Both passes over a 3x2x3 mesh (144 qps, three wavefronts pulling from the queue at once)."""
    out = _run(emu_env, "pass", "fcc", 2e-3, 3, 9, *([kin] if kin else []))
    assert out["flag1"]["nsing"] > 0  # the in-place pseudo-inverse branch is exercised
    for flag in ("flag0", "flag1"):
        for k, v in out[flag]["err"].items():
            # slip and stress (via slip) carry the pseudo-inverse's conditioning: 1e-12 observed
            assert v < (1e-9 if k in ("slip", "P", "T") else 1e-10), (flag, k, v)


def test_coop_passes_on_golden_case_states(emu_env):
    """fcc/simpleTension 8^3, the oracle's state after 6 increments (past yield, singular local systems)."""
    out = _run(emu_env, "golden", "fcc_simpleTension", 6)
    assert out["flag1"]["nsing"] > 0
    for flag in ("flag0", "flag1"):
        for k, v in out[flag]["err"].items():
            assert v < 1e-10, (flag, k, v)


def test_coop_with_twin_systems_equals_one_lane_routine(emu_env):
    """hcp/simpleTension's 18 slip + 6 twin systems (one-sided twin shear): the cooperative form
    equals the one-lane routine bit for bit; both equal the oracle except on the qps whose local
    problem meets the exactly singular twin system, where round-off picks the twin pair
    (tests/test_gpu_twinning.py)."""
    out = _run(emu_env, "twin", 256, 17)
    for flag in ("flag0", "flag1"):
        for k, v in out[flag]["coop_vs_lane"].items():
            assert v == 0.0, (flag, k, v)
        assert out[flag]["twin_qps"] > 60
        assert out[flag]["coop_vs_oracle_bad"] == out[flag]["lane_vs_oracle_bad"]
        assert len(out[flag]["lane_vs_oracle_bad"]) <= 8

