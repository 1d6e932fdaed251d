"""The reference's own model.cuh files compile unchanged against the engine's
CCalcModel contract (template/model.cuh:21-57; src/mcmcjob.cuh:30-87 for the
fields a model may read).

Each project's model.cuh is compiled for gfx950 into the generic path
(csrc/generic.hip, which includes "model.cuh" from the include path) straight
from the reference checkout -- nothing is copied.  Skipped where the reference
checkout or hipcc is absent.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from cuda_mcmc_amd import build

REF = "/root/reference"
PROJECTS = ["examples/search_Rosenbrock", "examples/solvsys", "examples/sgra_model_fit", "template"]


def _compile(project):
    d = os.path.join(REF, project)
    cmd = [build.HIPCC] + build.FLAGS + ["-I" + d, '-DMHRE_MODEL_NAME="%s/model.cuh"' % project,
                                         "--cuda-device-only", "-c", "-o", os.devnull,
                                         os.path.join(build.CSRC, "generic.hip")]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return project, r.returncode, r.stdout


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc not installed")
def test_reference_model_cuh_files_compile_unchanged():
    present = [p for p in PROJECTS if os.path.isfile(os.path.join(REF, p, "model.cuh"))]
    assert len(present) == 4
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(_compile, present))
    for project, rc, out in results:
        assert rc == 0, "%s/model.cuh does not compile:\n%s" % (project, out[-3000:])
