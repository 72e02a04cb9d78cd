"""The provider's host policies around the solver (crates/providers/src/ks.rs, pseudo.rs), exercised without a GPU:
the GPU tolerance floor through the library, the R15 retry policy by compiling its header against a scripted solve
(as the reference tests retry_cold_on_stall without a solver)."""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "atomic-sandbox_amd"))
import lira  # noqa: E402


def test_effective_gpu_tol_is_twice_the_f32_density_floor():
    # pseudo.rs:423-440: floor 5e-8 per valence electron, the tolerance raised to twice it and never lowered
    assert lira.effective_gpu_tol(1e-9, 8) == pytest.approx(8e-7)
    assert lira.effective_gpu_tol(1e-6, 8) == 1e-6
    assert lira.effective_gpu_tol(1e-9, 0) == 1e-9
    assert lira.effective_gpu_tol(1e-9, 12) == pytest.approx(1.2e-6)


POLICY_TEST = r"""
#include "provider_policy.hpp"
#include <cstdio>
#include <vector>
enum S { OK, NOT_CONVERGED, STALLED, NUMERICAL };
int main() {
    int bad = 0;
    auto run = [&](bool warm, std::vector<S> outcomes, S want, bool want_retry, std::vector<bool> want_calls) {
        std::vector<bool> calls;
        size_t i = 0;
        bool retried;
        S first;
        S got = lira::retry_cold_on_stall<S>(warm, NOT_CONVERGED, STALLED, [&](bool w) { calls.push_back(w); return outcomes[i++]; },
                                           &retried, &first);
        if (got != want || retried != want_retry || calls != want_calls) { bad++; std::printf("case %d failed\n", bad); }
    };
    run(true, {OK}, OK, false, {true});                                   // a warm success is kept
    run(true, {NOT_CONVERGED, OK}, OK, true, {true, false});              // warm cap reached: one cold retry
    run(true, {STALLED, OK}, OK, true, {true, false});                    // warm stall: one cold retry
    run(true, {STALLED, NOT_CONVERGED}, NOT_CONVERGED, true, {true, false});  // the cold outcome is returned
    run(false, {NOT_CONVERGED}, NOT_CONVERGED, false, {false});           // a cold failure is final
    run(true, {NUMERICAL}, NUMERICAL, false, {true});                     // other errors are not retried
    std::printf("%s\n", bad ? "FAIL" : "OK");
    return bad;
}
"""


def test_retry_cold_on_stall_policy():
    # ks.rs:585-601 (R15): only a warm attempt that ran out of steps or stalled is retried, exactly once, cold
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "policy.cpp")
        exe = os.path.join(d, "policy")
        with open(src, "w") as f:
            f.write(POLICY_TEST)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "atomic-sandbox_amd", "csrc"), src, "-o",
                        exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout


DEVICE_TEST = r"""
#include "provider_policy.hpp"
#include <cstdio>
#include <vector>
enum S { OK, DEVICE, CANCELLED, NUMERICAL };
int main() {
    int bad = 0;
    auto run = [&](std::vector<S> outcomes, std::vector<bool> pauses, S want, int want_failed, std::vector<int> want_ms) {
        size_t i = 0, j = 0;
        std::vector<int> ms;
        int failed = -1;
        S got = lira::retry_on_device_fault<S>(DEVICE, CANCELLED, [&] { return outcomes[i++]; },
                                              [&](int m) { ms.push_back(m); return pauses[j++]; }, &failed);
        if (got != want || failed != want_failed || ms != want_ms) { bad++; std::printf("case %d failed\n", bad); }
    };
    run({OK}, {}, OK, 0, {});                                              // a success is kept
    run({DEVICE, OK}, {true}, OK, 1, {100});                               // one failure: repeated after 100 ms
    run({DEVICE, DEVICE, OK}, {true, true}, OK, 2, {100, 400});            // two: 100 then 400 ms
    run({DEVICE, DEVICE, DEVICE}, {true, true}, DEVICE, 2, {100, 400});    // the third failure is final
    run({DEVICE, OK}, {false}, CANCELLED, 0, {100});                       // cancelled during the pause
    run({NUMERICAL}, {}, NUMERICAL, 0, {});                                // other errors are not retried
    // witness_cap_for: the largest w in (3, 32] whose footprint fits, never below 3
    auto fp = [](int w) { return 100ull + 10ull * w; };
    if (lira::witness_cap_for_budget(fp, 0, 3, 32) != 3) bad++;
    if (lira::witness_cap_for_budget(fp, 100 + 10 * 7, 3, 32) != 7) bad++;
    if (lira::witness_cap_for_budget(fp, 100 + 10 * 7 + 9, 3, 32) != 7) bad++;
    if (lira::witness_cap_for_budget(fp, 1ull << 40, 3, 32) != 32) bad++;
    std::printf("%s\n", bad ? "FAIL" : "OK");
    return bad;
}
"""


def test_device_retry_and_witness_budget_policies():
    # ks.rs:317-350 (the "GPU busy" retry: 100 ms, 400 ms, then the error) and ks.rs:400-428 (the budget's witness cap)
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "device.cpp")
        exe = os.path.join(d, "device")
        with open(src, "w") as f:
            f.write(DEVICE_TEST)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "atomic-sandbox_amd", "csrc"), src, "-o",
                        exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.strip() == "OK", out.stdout


def test_solve_footprint_and_budget_cap():
    # the footprint grows with the witness count and the precision; the budget's cap is the largest that fits
    f = [lira.solve_footprint_bytes(95, lira.MIXED, 5, 5, w) for w in range(3, 33)]
    assert all(b > a for a, b in zip(f, f[1:]))
    assert lira.solve_footprint_bytes(95, lira.FP64, 5, 5, 3) > lira.solve_footprint_bytes(95, lira.MIXED, 5, 5, 3)
    assert lira.witness_cap_for_budget(95, lira.MIXED, 5, 5, 1) == 3
    assert lira.witness_cap_for_budget(95, lira.MIXED, 5, 5, 1 << 50) == 32
    budget = lira.solve_footprint_bytes(95, lira.MIXED, 5, 5, 9) + 1
    assert lira.witness_cap_for_budget(95, lira.MIXED, 5, 5, budget) == 9
    assert lira.solve_footprint_bytes(2, lira.MIXED, 1, 1, 3) == 0
