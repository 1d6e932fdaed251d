"""The division-free step-size adaptation of the quadratic-system step kernel
against the reference's divisions (src/gpu_mcmc.cu:493, :518-522), on the
host: tests/step_adapt_check.hip runs the product's own helpers
(cuda_mcmc_amd/csrc/mhre_device.h) and checks

* step_shrink(sd) == (float)((double)sd / 1.01) for all 2^32 float bit patterns;
* rate_below(sh, den, rate_threshold(T)) == ((float)sh / (float)den < T) over
  sh in [-3, 2100], den in [0, 2048] and 14 targets;
* mod_small(cnt, n) == cnt % n for cnt below 2^23 and 19 moduli.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_build", "step_adapt_check")


def test_adaptation_forms_match_the_divisions():
    if not os.path.isfile(EXE):
        pytest.fail("tests/_build/step_adapt_check is not built (python -c 'import __graft_entry__ as g; g.build()')")
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    lines = {l.split()[0]: l.split() for l in r.stdout.splitlines() if "mismatches" in l}
    assert set(lines) == {"step_shrink", "rate_below", "mod_small"}, r.stdout
    assert int(lines["step_shrink"][2]) == 1 << 32
    for name, f in lines.items():
        assert int(f[4]) == 0, r.stdout
    assert r.returncode == 0, r.stdout
