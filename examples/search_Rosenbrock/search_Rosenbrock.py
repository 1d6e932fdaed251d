"""Global minimum of the Rosenbrock function with MH-RE (Python 3 counterpart
of the reference's examples/search_Rosenbrock/search_Rosenbrock.py, making the
same imgpu.Mcgpu calls).

    f(x, y) = (1 - x)^2 + 100 (y - x^2)^2,  minimum 0 at (1, 1)

The library with this directory's model.cuh is built by
``python -m cuda_mcmc_amd.build --model-dir examples/search_Rosenbrock``.
Pass --plot for the sample histograms (matplotlib).
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
os.environ.setdefault("MHRE_LIB", os.path.join(HERE, "libmhre.so"))

import numpy as np  # noqa: E402
from numpy import array, float32, int32, zeros  # noqa: E402

from cuda_mcmc_amd import imgpu  # noqa: E402

pdescr1 = array((1, 1), dtype=int32)
pmint1 = array((-10., -10.), dtype=float32)
pmaxt1 = array((10., 10.), dtype=float32)
coor1 = array((2, 2, 0.25, 0.5), dtype=float32)   # unused by this model
dat1 = zeros(1, dtype=float32)                    # f is compared with zero

rosen = imgpu.Mcgpu(coor=coor1, dat=dat1, pdescr=pdescr1, pmint=pmint1, pmaxt=pmaxt1,
                    nbeta=32, nseq=14, nburn=400, niter=1000)
rosen.burnin_and_search()

chi2 = rosen.chi2
im = chi2.argmin()
pout = rosen.pout
print("The minimum chi^2 at chi2[%d] = %f" % (im, chi2[im]))
print("pout[:, %d] = " % im, pout[:, im])

imin = np.where(chi2 < 0.00001)[0]
order = np.argsort(chi2[imin])
print("\nExact solution: (x, y) = (%10f, %10f)\n" % (1, 1))
print("Best minima found:\n     x           y       chi^2")
for i in imin[order][:20]:
    print("%10f %10f %10f" % (pout[0, i], pout[1, i], chi2[i]))

if "--plot" in sys.argv:
    import matplotlib.pyplot as plt
    plt.figure(figsize=(12, 6))
    plt.subplot(121); plt.hist(pout[0, :], 50, color="b"); plt.grid(1); plt.xlabel("x")
    plt.title("MCMC Output Distribution for X")
    plt.subplot(122); plt.hist(pout[1, :], 50, color="b"); plt.grid(1); plt.xlabel("y")
    plt.title("MCMC Output Distribution for Y")
    plt.show()
