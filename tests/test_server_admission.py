"""Admission rule of the serving loop (sdot_amd/csrc/admission.h), host-only: a
small C++ driver compiled with g++ replays the FIFO admission of one lane under
the rule, with the 3D workers' idle reservations counted as resident."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include <deque>
#include "admission.h"
using namespace sdot_amd;
int main() {
    int fails = 0;
    auto check = [&](bool c, const char *what) { if (!c) { std::printf("FAIL %s\n", what); ++fails; } };
    const double budget = 10.0, reserve = 1.6;
    // two 3D workers have each kept their handle reservation; nothing runs
    double resident = 2 * reserve;
    long running = 0;
    // a 2D request that fits the budget but not beside the idle reservations starts
    check(srv_admit(resident, 8.0, budget, running), "idle device admits a request larger than budget - reservations");
    // ... but not while something runs
    check(!srv_admit(resident + 1.0, 8.0, budget, 1), "a running request keeps it queued");
    check(srv_admit(resident, 6.0, budget, 1), "a request that fits beside them starts");
    // requests that could not fit even alone are rejected at submission
    check(srv_never_fits(11.0, 0.0, budget), "2D request above the budget");
    check(srv_never_fits(9.0, reserve, budget), "3D request whose worker handle would not fit");
    check(!srv_never_fits(8.0, 0.0, budget), "a request within the budget is accepted");
    // FIFO replay: requests of 8, 3, 3 with the reservations resident; the first starts on
    // the idle device, the others wait for it, then both start together
    std::deque<double> q = {8.0, 3.0, 3.0};
    int started = 0;
    auto admit_all = [&] {
        while (!q.empty() && srv_admit(resident, q.front(), budget, running)) {
            resident += q.front();
            ++running;
            ++started;
            q.pop_front();
        }
    };
    admit_all();
    check(started == 1 && running == 1, "only the head starts");
    resident -= 8.0;   // it finishes
    --running;
    admit_all();
    check(started == 3 && q.empty(), "the queue drains once the head has finished");
    std::printf(fails ? "admission: %d failures\n" : "admission: ok\n", fails);
    return fails ? 1 : 0;
}
"""


def test_admission_rule(tmp_path):
    src = tmp_path / "admission_driver.cpp"
    exe = tmp_path / "admission_driver"
    src.write_text(DRIVER)
    inc = os.path.join(ROOT, "sdot_amd", "csrc")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", inc, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "admission: ok" in r.stdout
