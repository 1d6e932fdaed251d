"""The hardware sine and cosine behind sincosrev (argument in revolutions,
reduced by h - rint(h)) and sincosrev_hw (no reduction, |h| <= 255) against
float64: the phases of the ring and Gaussian terms
use them in the fused and gradient kernels, so their error enters every model
visibility.  The bound 2e-7 keeps the model within the chi^2 tolerance of
1e-5 relative; the reduction keeps it at any argument size."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_build", "sincos_check")

pytestmark = pytest.mark.gpu


def test_hardware_sincos_accuracy():
    if not os.path.isfile(EXE):
        pytest.fail("tests/_build/sincos_check is not built (python -c 'import __graft_entry__ as g; g.build()')")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120, check=True).stdout
    print(out)
    rows = [line.split() for line in out.splitlines() if line.startswith("range ")]
    assert len(rows) == 4
    for r in rows:
        es, ec = float(r[3]), float(r[5])
        # the reduction h - rint(h) is exact, so the error does not grow
        # with the argument
        assert es < 2e-7 and ec < 2e-7, r
    # without the reduction, inside the instructions' input domain (the step
    # kernel's form for fits whose phases the host bounds by 64 revolutions)
    rows = [line.split() for line in out.splitlines() if line.startswith("range_hw ")]
    assert len(rows) == 4
    for r in rows:
        es, ec = float(r[3]), float(r[5])
        assert es < 2e-7 and ec < 2e-7, r
