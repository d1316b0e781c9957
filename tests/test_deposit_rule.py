"""The product's bucket-deposit rule (teufel_amd/csrc/tg_deposit.h, the code the
HIP kernels run) compiled for the host and compared with the oracle's
restatement of src/particle.cpp:788-865 on 200k random steps and window
splits, and the per-lane walk's bucket sums (the lane accumulator) on 20k
random whole trajectories through windows with and without the spill bucket
(tests/cpp/deposit_rule.cpp).  CPU only."""
import os
import subprocess

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_deposit_rule_matches_oracle(tmp_path):
    oracle.build()
    exe = tmp_path / "deposit_rule"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas",
                    "-o", str(exe), os.path.join(HERE, "cpp", "deposit_rule.cpp"), oracle.LIB_PATH,
                    "-Wl,-rpath," + os.path.dirname(oracle.LIB_PATH)], check=True)
    r = subprocess.run([str(exe), "200000"], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
