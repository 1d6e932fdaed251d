"""CPU oracle for the MH-RE hot path -- TEST INFRASTRUCTURE ONLY.

This package is a float64 NumPy restatement of the reference algorithm
(MITHaystack/cuda_mcmc, ``src/gpu_mcmc.cu:228-763`` and
``examples/sgra_model_fit/model.cuh:9-462``) used as the correctness checker
for the HIP engine in ``cuda_mcmc_amd``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline -- never as a
code path of the product.

Modules
-------
philox    Philox4x32-10 (Salmon et al. 2011) and the engine's counter layout.
          Pinned by the Random123 known-answer vectors.
models    fp64 ring+Gaussian visibility models (9- and 13-parameter) and the
          toy models (Rosenbrock, solvsys, quadratic systems).  Pinned by
          closed forms (zero spacing, Airy disk, pure Gaussian), Hermitian
          symmetry, linearity in Zsp, a direct image-plane Fourier sum of the
          slashed ring, and a point-by-point restatement (in the tests) of
          the reference's second fp64 mirror ``src/obsdatafit.py:1476-1608``;
          the quadratic systems by the reference's solvsys roots, a
          hand-computed case and the antisymmetric-part invariant.
chi2      closures and chi^2 terms (``model.cuh:73-111``,
          ``gpu_mcmc.cu:305-323``).  Pinned by hand-computed 3-station cases.
mhre      the MH-RE update rule (``gpu_mcmc.cu:245-626``) driven by the same
          Philox streams as the device.  Pinned by decision-level checks
          against brute force on tiny chains and statistical checks against
          targets with known moments.
uvdata    restatement of the uv loader and closure builder
          (``src/imgpu.py:1229-1325, 1375-1474``).  Pinned by independent
          counts on the fixture files.
synth     synthetic data generators for the benchmark configurations.

Parity status of every function is recorded in each module's docstring.
"""
