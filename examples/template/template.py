"""Template project: the model.cuh in this directory returns a constant, so
the sampler just explores the prior box (Python 3 counterpart of the
reference's template/template.py, making the same imgpu.Mcgpu calls).

Copy this directory, put your model() into model.cuh and build it with
``python -m cuda_mcmc_amd.build --model-dir <your project>``.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
os.environ.setdefault("MHRE_LIB", os.path.join(HERE, "libmhre.so"))

from numpy import array, float32, int32  # noqa: E402

from cuda_mcmc_amd import imgpu  # noqa: E402

pdescr1 = array((1, 1), dtype=int32)
pmint1 = array((-5., -5.), dtype=float32)
pmaxt1 = array((5., 5.), dtype=float32)

solver = imgpu.Mcgpu(pdescr=pdescr1, pmint=pmint1, pmaxt=pmaxt1)
solver.burnin_and_search()

chi2 = solver.chi2
im = chi2.argmin()
pout = solver.pout
print("The minimum chi^2 at chi2[%d] = %f" % (im, chi2[im]))
print("pout[:, %d] = " % im, pout[:, im])
print("sample mean (prior centre expected):", pout.mean(axis=1))
