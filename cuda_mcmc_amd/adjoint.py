"""chi^2 gradients, Gauss-Newton polish and Laplace covariance of the
9-parameter ring+Gaussian model (``imodel`` 1).

The reference samples without derivatives; this module adds them on the same
data arrays that ``calc_modchi2`` takes (``src/mcmc.h:23-27``), so that
``chi2(P)`` here equals the ``chi2c`` of ``calc_modchi2`` for the same
parameter rows.  The work runs in the HIP kernels behind the C ABI
(``mhre_eht_chi2_grad``, ``mhre_eht_gn_system``); torch supplies device
memory and the stream only.

EhtData            one data set packed on the current GPU.
chi2(P, data)      differentiable batched chi^2 (torch.autograd.Function):
                   P [n, nptot] float32 on the GPU -> chi^2 [n] float64.
gauss_newton(...)  Levenberg-Marquardt polish of parameter rows inside the
                   prior box (the 9 x 9 solves run on the host in float64).
laplace_covariance(data, P, ivar)
                   (J^T J)^-1 over the varied parameters: the covariance of
                   the Gaussian approximation of exp(-chi^2/2) at P.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import EngineError

NP9 = 9
_IU = np.triu_indices(NP9)


def _stream(t):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _dptr(t):
    return C.c_void_p(t.data_ptr())


class EhtData:
    """The calc_modchi2 data arrays of one observation, packed on the GPU.

    coor   interleaved (u, v) in Glambda [+ (u1, v1, u2, v2) per closure when
           idat[2] = cpExt = 1], dat / std2r the amplitudes then the closure
           phases and their 1/sigma^2, icoor the closure triplets, idat =
           [nvis, ndatm, cpExt].  Same layout as ``Mcgpu_Sgra`` builds
           (``src/imgpu.py:759-831``).
    """

    def __init__(self, coor, dat, std2r, icoor, idat, nptot=NP9, imodel=1, lib=None):
        if not torch.cuda.is_available():
            raise EngineError("EhtData needs a GPU")
        self.lib = _lib.load(lib)
        coor = np.ascontiguousarray(coor, np.float32).ravel()
        dat = np.ascontiguousarray(dat, np.float32).ravel()
        std2r = np.ascontiguousarray(std2r, np.float32).ravel()
        icoor = np.ascontiguousarray(icoor if icoor is not None else np.zeros(0), np.int32).ravel()
        idat = np.ascontiguousarray(idat, np.int32).ravel()
        self.nptot = int(nptot)
        self.device = torch.device("cuda", torch.cuda.current_device())
        h = C.c_void_p()
        rc = self.lib.mhre_eht_data_create(_lib.ptr(coor, C.c_float), _lib.ptr(dat, C.c_float),
                                           _lib.ptr(std2r, C.c_float), _lib.ptr(icoor, C.c_int),
                                           _lib.ptr(idat, C.c_int), int(imodel), coor.size, dat.size,
                                           icoor.size, idat.size, self.nptot, C.byref(h))
        _lib.check(self.lib, rc, "mhre_eht_data_create")
        self.handle = h

    @classmethod
    def from_dataset(cls, d, **kw):
        """From a dict with coor, dat, std2r, icoor, nvis, ncph (synthdata layout)."""
        idat = np.array([d["nvis"], 0, int(d.get("cpext", 0))], np.int32)
        return cls(d["coor"], d["dat"], d["std2r"], d["icoor"], idat, **kw)

    def info(self):
        """How the data were packed: dict(ring_gradient, rounds, cgroup, ring_ok)."""
        out = (C.c_int * 4)()
        _lib.check(self.lib, self.lib.mhre_eht_data_info(self.handle, out, 4), "mhre_eht_data_info")
        return dict(ring_gradient=bool(out[0]), rounds=int(out[1]), cgroup=int(out[2]), ring_ok=bool(out[3]))

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.mhre_eht_data_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rows(self, P):
        if not isinstance(P, torch.Tensor) or P.device.type != "cuda":
            raise TypeError("parameters must be a CUDA tensor")
        if P.dtype != torch.float32 or P.dim() != 2 or not P.is_contiguous():
            raise TypeError("parameters must be a contiguous float32 [n, nptot] tensor")
        if P.shape[1] < NP9 or P.shape[1] > 32:
            raise ValueError("parameter rows have %d entries, need 9..32" % P.shape[1])
        return P.shape[0], P.shape[1]

    def chi2_grad(self, P, need_grad=True):
        """chi^2 [n] (float64) and, if need_grad, d chi^2/dP [n, nptot] (float32)."""
        n, ldp = self._rows(P)
        chi = torch.empty(n, dtype=torch.float64, device=P.device)
        grad = torch.empty((n, ldp), dtype=torch.float32, device=P.device) if need_grad else None
        rc = self.lib.mhre_eht_chi2_grad(self.handle, _dptr(P), n, ldp, _dptr(chi),
                                         _dptr(grad) if grad is not None else None, _stream(P))
        _lib.check(self.lib, rc, "mhre_eht_chi2_grad")
        return chi, grad

    def gn_system(self, P):
        """chi^2 [n], J^T r [n, 9] and J^T J [n, 9, 9] of the weighted residuals."""
        n, ldp = self._rows(P)
        chi = torch.empty(n, dtype=torch.float64, device=P.device)
        jtr = torch.empty((n, NP9), dtype=torch.float64, device=P.device)
        tri = torch.empty((n, NP9 * (NP9 + 1) // 2), dtype=torch.float64, device=P.device)
        rc = self.lib.mhre_eht_gn_system(self.handle, _dptr(P), n, ldp, _dptr(chi), _dptr(jtr), _dptr(tri),
                                         _stream(P))
        _lib.check(self.lib, rc, "mhre_eht_gn_system")
        jtj = torch.zeros((n, NP9, NP9), dtype=torch.float64, device=P.device)
        iu = torch.as_tensor(np.stack(_IU), device=P.device)
        jtj[:, iu[0], iu[1]] = tri
        jtj = jtj + jtj.transpose(1, 2) - torch.diag_embed(torch.diagonal(jtj, dim1=1, dim2=2))
        return chi, jtr, jtj


class _Chi2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P, data):
        chi, grad = data.chi2_grad(P, need_grad=ctx.needs_input_grad[0])
        ctx.save_for_backward(grad if grad is not None else torch.empty(0, device=P.device))
        return chi

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (g.to(grad.dtype)[:, None] * grad, None)


def chi2(P, data):
    """Differentiable chi^2 of each parameter row of P against data."""
    return _Chi2.apply(P, data)


def _clip_wrap(p, pmin, pmax, pdescr):
    ang = pdescr == 2
    p[:, ang] = p[:, ang] - 2.0 * np.pi * np.rint(p[:, ang] / (2.0 * np.pi))
    return np.clip(p, pmin, pmax)


def gauss_newton(data, P0, pmin, pmax, pdescr, iters=20, lam=1e-3, rtol=1e-12):
    """Levenberg-Marquardt polish of the rows of P0 (numpy or tensor [n, nptot]).

    Only parameters with pdescr != 0 move; steps are clipped into
    [pmin, pmax] (angles, pdescr == 2, wrapped first).  Each iteration solves
    (J^T J + lam diag J^T J) d = -J^T r per row and keeps the step if chi^2
    drops (lam / 10), else rejects it (lam * 10).  Returns (P, chi^2) as
    numpy arrays (float32, float64).
    """
    pmin = np.asarray(pmin, np.float64)
    pmax = np.asarray(pmax, np.float64)
    pdescr = np.asarray(pdescr)
    var = np.nonzero(pdescr[:NP9] != 0)[0]
    P = (P0.detach().cpu().numpy() if isinstance(P0, torch.Tensor) else np.asarray(P0)).astype(np.float64)
    n = P.shape[0]
    lamv = np.full(n, float(lam))
    dev = data.device
    Pt = torch.as_tensor(P.astype(np.float32), device=dev).contiguous()
    chi, jtr, jtj = data.gn_system(Pt)
    chi, jtr, jtj = chi.cpu().numpy(), jtr.cpu().numpy(), jtj.cpu().numpy()
    active = np.ones(n, bool)
    for _ in range(iters):
        if not active.any():
            break
        step = np.zeros_like(P)
        for r in np.nonzero(active)[0]:
            A = jtj[r][np.ix_(var, var)]
            A = A + lamv[r] * np.diag(np.diag(A))
            try:
                step[r, var] = -np.linalg.solve(A, jtr[r, var])
            except np.linalg.LinAlgError:
                active[r] = False
        trial = _clip_wrap(P + step, pmin, pmax, pdescr).astype(np.float32)
        Tt = torch.as_tensor(trial, device=dev).contiguous()
        c2, j2, jj2 = data.gn_system(Tt)
        c2, j2, jj2 = c2.cpu().numpy(), j2.cpu().numpy(), jj2.cpu().numpy()
        better = active & (c2 < chi)
        conv = better & ((chi - c2) <= rtol * np.maximum(chi, 1e-300))
        P[better] = trial[better]
        chi[better], jtr[better], jtj[better] = c2[better], j2[better], jj2[better]
        lamv[better] /= 10.0
        lamv[active & ~better] *= 10.0
        active &= ~conv & (lamv < 1e12)
    return P.astype(np.float32), chi


def laplace_covariance(data, P, pdescr):
    """Covariance (J^T J)^-1 of the varied parameters at each row of P.

    Returns (cov [n, nv, nv], var) with var the indices of the varied
    parameters.  For exp(-chi^2/2) the Hessian of chi^2/2 is J^T J in the
    Gauss-Newton approximation.
    """
    pdescr = np.asarray(pdescr)
    var = np.nonzero(pdescr[:NP9] != 0)[0]
    Pn = P.detach().cpu().numpy() if isinstance(P, torch.Tensor) else np.asarray(P)
    Pt = torch.as_tensor(np.ascontiguousarray(Pn, np.float32), device=data.device)
    _, _, jtj = data.gn_system(Pt)
    jtj = jtj.cpu().numpy()
    cov = np.stack([np.linalg.inv(m[np.ix_(var, var)]) for m in jtj])
    return cov, var
