"""The Python 3 example drivers run end to end on their own project libraries
and find the known answers: the Rosenbrock minimum (1, 1) and both roots of the
solvsys system (examples/*/).  Each script runs in a child process, as a user
would start it.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(rel):
    env = dict(os.environ)
    env.pop("MHRE_LIB", None)
    p = subprocess.run([sys.executable, os.path.join(ROOT, rel)], capture_output=True, text=True, timeout=110,
                       env=env, cwd=ROOT)
    print(p.stdout[-2000:])
    assert p.returncode == 0, p.stderr[-2000:]
    return p.stdout


def _pairs(out):
    return np.array([[float(a), float(b)] for a, b, _ in
                     re.findall(r"^\s*(-?\d+\.\d+)\s+(-?\d+\.\d+)\s+(\d+\.\d+)\s*$", out, re.M)])


def test_rosenbrock_example():
    best = _pairs(_run("examples/search_Rosenbrock/search_Rosenbrock.py"))
    assert len(best) > 0
    np.testing.assert_allclose(best[0], [1.0, 1.0], atol=0.02)


def test_solvsys_example():
    best = _pairs(_run("examples/solvsys/solvsys.py"))
    rx1, rx2 = (-4. - np.sqrt(832.)) / 34., (-4. + np.sqrt(832.)) / 34.
    roots = np.array([[rx1, np.sqrt(1 - rx1 ** 2)], [rx2, np.sqrt(1 - rx2 ** 2)]])
    d = np.abs(best[:, None, :] - roots[None]).max(axis=2)
    assert np.all(d.min(axis=1) < 0.02)
    assert np.all((d < 0.02).any(axis=0)), "both roots found"


def test_solvsys_api_and_template_examples():
    out = _run("examples/solvsys/solvsys_api.py")
    assert "minimum chi^2" in out
    out = _run("examples/template/template.py")
    assert "sample mean" in out
