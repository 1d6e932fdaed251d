"""Roots of the system
       x^2 + y^2 = 1
    -0.25 x + y = 0.5
with MH-RE (Python 3 counterpart of the reference's examples/solvsys/solvsys.py,
making the same imgpu.Mcgpu calls).  Exact roots: (-0.966, 0.2585) and
(0.7307, 0.6827).

The library with this directory's model.cuh is built by
``python -m cuda_mcmc_amd.build --model-dir examples/solvsys``.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
os.environ.setdefault("MHRE_LIB", os.path.join(HERE, "libmhre.so"))

import numpy as np  # noqa: E402
from numpy import array, float32, int32, sqrt  # noqa: E402

from cuda_mcmc_amd import imgpu  # noqa: E402

pdescr1 = array((1, 1), dtype=int32)
pmint1 = array((-5., -5.), dtype=float32)
pmaxt1 = array((5., 5.), dtype=float32)
coor1 = array((2, 2, 0.25, 0.5), dtype=float32)   # x^coor0 + y^coor1 ; -coor2 x + y
dat1 = array((1, 0.5), dtype=float32)

solver = imgpu.Mcgpu(coor=coor1, dat=dat1, pdescr=pdescr1, pmint=pmint1, pmaxt=pmaxt1,
                     nbeta=32, nseq=14, nburn=400, niter=5000)
solver.burnin_and_search()

chi2 = solver.chi2
im = chi2.argmin()
pout = solver.pout
print("The minimum chi^2 at chi2[%d] = %f" % (im, chi2[im]))
print("pout[:, %d] = " % im, pout[:, im])

rx1 = (-4. - sqrt(832.)) / 34.
rx2 = (-4. + sqrt(832.)) / 34.
ry1 = sqrt(1. - rx1 ** 2)
ry2 = sqrt(1. - rx2 ** 2)
print("\nExact roots: (%.4f, %.4f), (%.4f, %.4f)" % (rx1, ry1, rx2, ry2))
imin = np.where(chi2 < 0.0002)[0]
order = np.argsort(chi2[imin])
print("Best root pairs found:\n     x           y       chi^2")
for i in imin[order][:20]:
    print("%10f %10f %10f" % (pout[0, i], pout[1, i], chi2[i]))
