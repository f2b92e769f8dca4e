"""Device-free decisions of the library, checked on the host (tests/host_checks.cpp):

* the event-pair bookkeeping of the profiled kernel (csrc/ba_ledger.hpp): the count of kept
  launches is clamped to the pairs that exist, so a product counter read from device memory can
  never index past them;
* the PCG product form per LM iteration (csrc/ba_policy.hpp): 9 against 10 PCG iterations, the
  first iteration and a failed solve, forced forms, no explicit system.
"""

import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parent


def test_host_checks(tmp_path):
    exe = tmp_path / "host_checks"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fsanitize=address,undefined",
                    "-o", str(exe), "-I", str(ROOT.parent / "include"), str(ROOT / "host_checks.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert out.strip() == "ok", out
