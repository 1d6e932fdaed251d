"""Compile-time guard: the hot kernels keep all state in registers and LDS.

A per-lane stack (scratch memory) in the fused step kernel turns into HBM
traffic and slows every iteration; this checks the compiler's resource
report for the EHT step, exchange and evaluation kernels (no GPU needed).
"""
import os
import re
import subprocess

import pytest

from cuda_mcmc_amd import build

SRC = os.path.join(build.CSRC, "engine.hip")


def _resources():
    cmd = [build.HIPCC] + build.FLAGS + ["--cuda-device-only", "-c", "-o", os.devnull, SRC,
                                         "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True).stdout
    res, name = {}, None
    for line in out.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            res[name] = {}
        m = re.search(r"(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and name:
            res[name][m.group(1).split()[0]] = int(m.group(2))
    return res


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc not installed")
def test_hot_kernels_use_no_scratch():
    res = _resources()
    hot = [n for n in res if "k_step_eht" in n or "k_exchange" in n or "k_eval_eht" in n]
    assert len(hot) >= 10
    for n in hot:
        assert res[n].get("ScratchSize", 0) == 0, (n, res[n])
    grid = [n for n in res if "k_step_ehtINS_5Sgra9ELi4ELb0" in n]
    assert grid and res[grid[0]]["Occupancy"] >= 4, res[grid[0]]
