"""float64 restatement of the reference's MH-RE update rule (oracle side).

Follows ``src/gpu_mcmc.cu`` kernel by kernel, fed by the engine's Philox
streams (``oracle/philox.py``) so that the device and this oracle consume the
same numbers:

init_pcur       ``gpu_mcmc.cu:245-267``   uniform start inside the prior box,
                                          angle wrap, flag = 1.
initial chi^2   ``gpu_mcmc.cu:714-719``   chi^2 of every chain.
mh_substep      ``gpu_mcmc.cu:361-474``   random coordinate, Gaussian step,
                                          wrap, prior check, accept test,
                                          revert on reject, n_acpt++.
adjust_steps    ``gpu_mcmc.cu:480-526``   nadj-deep accept history ring;
                                          rate < target -> pstp/1.01 else
                                          pstp*1.01; n_cnt++ (burn-in only).
exchange        ``gpu_mcmc.cu:533-626``   nbeta random adjacent-pair swap
                                          attempts per sequence; n_exch++ and
                                          pout/tout/chi2 recording in the main
                                          stage.
run             ``gpu_mcmc.cu:682-763``   burn-in then sampling.

Arithmetic is float64.  State arrays keep the reference's storage types
(float32 parameters, steps and chi^2; int32 counters) so each update rounds on
store exactly as the reference's ``float`` arrays do; pass
``state_dtype=np.float64`` for a pure-fp64 run.

Decisions are taken as log(u) <= -beta/2 * dchi2 (equivalent to the
reference's u <= min(exp(.), 1)).  Every decision's margin log(u) - x is
recorded so tests can exclude near-ties.

Deliberate differences from the reference (each is a documented quirk):
* one exchange attempt loop is skipped when nbeta == 1 (the reference's
  ``do {} while (k >= nbeta-1)`` never terminates there);
* the angle wrap is p - 2*pi*rint(p / 2*pi) in float64 (same principal value
  as atan2(sin p, cos p) away from exactly +-pi).

Parity: pinned by brute-force single-chain checks of each rule, by
statistical checks on targets with known moments (1-D/2-D Gaussians) and by
the Rosenbrock / solvsys known answers (tests/test_oracle_mhre.py).
"""
import numpy as np
from . import philox as px

TWO_PI = 2.0 * np.pi


def wrap_angle(p):
    p = np.asarray(p, dtype=np.float64)
    return p - TWO_PI * np.rint(p / TWO_PI)


class Problem:
    """Static description of one fit (the reference's input arrays).

    chi2_fn maps a float64 parameter stack [n, nptot] to chi^2 [n].
    """

    def __init__(self, chi2_fn, pdescr, pmint, pmaxt, beta, nseq, ptotal=None,
                 nadj=100, seed=0, adapt_target=0.30):
        self.chi2_fn = chi2_fn
        self.pdescr = np.asarray(pdescr, dtype=np.int32)
        self.nptot = self.pdescr.size
        self.ivar = np.nonzero(self.pdescr)[0].astype(np.int32)
        self.nprm = self.ivar.size
        self.invar = np.full(self.nptot, -1, dtype=np.int32)
        self.invar[self.ivar] = np.arange(self.nprm, dtype=np.int32)
        self.pmint = np.asarray(pmint, dtype=np.float32)
        self.pmaxt = np.asarray(pmaxt, dtype=np.float32)
        self.beta = np.asarray(beta, dtype=np.float32)
        self.nbeta = self.beta.size
        self.nseq = int(nseq)
        self.ptotal = (np.zeros(self.nptot, np.float32) if ptotal is None
                       else np.asarray(ptotal, dtype=np.float32))
        self.nadj = int(nadj)
        self.seed = int(seed)
        self.key = px.key_of(seed)
        self.adapt_target = np.float32(adapt_target)

    @property
    def nchain(self):
        return self.nbeta * self.nseq


class State:
    """Chain state in the reference's layouts (src/mcmc_interf.pyx:232-263)."""

    def __init__(self, pb, niter=0, state_dtype=np.float32):
        B, S, N, P = pb.nbeta, pb.nseq, pb.nptot, pb.nprm
        fd = state_dtype
        self.fd = fd
        self.pcur = np.tile(pb.ptotal.astype(fd), (B, S, 1))
        pmid = (pb.pmint[pb.ivar].astype(np.float64) + pb.pmaxt[pb.ivar]) / 2.0
        self.pcur[:, :, pb.ivar] = pmid.astype(fd)
        self.chi2c = np.zeros((B, S), fd)
        step = (pb.pmaxt[pb.ivar].astype(np.float64) - pb.pmint[pb.ivar]) / 50.0
        self.pstp = np.broadcast_to(step.astype(fd)[:, None, None], (P, B, S)).copy()
        self.n_cnt = np.ones((P, B, S), np.int32)
        self.n_acpt = np.zeros((P, B, S), np.int32)
        self.n_hist = np.zeros((pb.nadj, P, B, S), np.int32)
        self.n_exch = np.zeros((B, S), np.int32)
        self.tcur = np.tile(np.arange(B, dtype=np.int32)[:, None], (1, S))
        self.flag = np.zeros((B, S), np.int32)
        self.ptent = np.zeros((B, S), fd)
        self.ptentn = np.zeros((B, S), np.int32)
        self.pout = np.zeros((P, B, S, niter), fd)
        self.tout = np.zeros((B, S, niter), np.int32)
        self.chi2 = np.zeros((S, niter), fd)

    def copy(self):
        new = State.__new__(State)
        new.__dict__ = {k: (v.copy() if isinstance(v, np.ndarray) else v)
                        for k, v in self.__dict__.items()}
        return new


def init_pcur(pb, st):
    """gpu_mcmc.cu:245-267 (uniform in prior) then the initial chi^2 (:714-719)."""
    B, S = pb.nbeta, pb.nseq
    chain = np.arange(B * S, dtype=np.uint64)
    pc = st.pcur.reshape(B * S, pb.nptot)
    for i in range(pb.nprm):
        k = pb.ivar[i]
        w = px.philox4x32(chain, px.INIT_ITER, i, px.INIT, *pb.key)[0]
        p = np.float64(pb.pmint[k]) + (np.float64(pb.pmaxt[k]) - np.float64(pb.pmint[k])) * px.u01(w)
        if pb.pdescr[k] == 2:
            p = wrap_angle(p)
        pc[:, k] = p.astype(st.fd)
    st.flag[:] = 1
    st.chi2c[:] = pb.chi2_fn(pc.astype(np.float64)).reshape(B, S).astype(st.fd)


def _pick(n, w0, retry_ctr):
    """pick with the reference's redraw loop; retry_ctr(r) -> words for retry r."""
    k = px.pick(w0, n)
    bad = k >= n
    r = 1
    while np.any(bad) and r <= px.MAX_RETRY:
        w = retry_ctr(r)
        k = np.where(bad, px.pick(w, n), k)
        bad = k >= n
        r += 1
    return np.where(bad, n - 1, k)


def mh_substep(pb, st, it, sub, burn, trace=None):
    """One gen_proposal + calc_chi2_terms + spmp_mcmc (+ adjust_steps) round."""
    B, S, N, P = pb.nbeta, pb.nseq, pb.nptot, pb.nprm
    C = B * S
    chain = np.arange(C, dtype=np.uint64)
    w0, w1, w2, w3 = px.philox4x32(chain, it, sub, 0, *pb.key)
    k = _pick(P, w0, lambda r: px.philox4x32(chain, it, sub, r, *pb.key)[0])
    t = pb.ivar[k]
    pc = st.pcur.reshape(C, N)
    pstp = st.pstp.reshape(P, C)
    chi2c = st.chi2c.reshape(C)
    flag = st.flag.reshape(C)
    old = pc[np.arange(C), t].copy()
    st.ptent.reshape(C)[:] = old
    st.ptentn.reshape(C)[:] = t
    sd = pstp[k, np.arange(C)].astype(np.float64)
    p = old.astype(np.float64) + sd * px.normal(w1, w2)
    ang = pb.pdescr[t] == 2
    p = np.where(ang, wrap_angle(p), p)
    pn = p.astype(st.fd)
    inprior = ~((pn < pb.pmint[t]) | (pn > pb.pmaxt[t]))
    idx = np.nonzero(inprior)[0]
    chi2 = np.full(C, np.nan)
    margin = np.full(C, np.inf)
    accept = np.zeros(C, bool)
    if idx.size:
        trial = pc[idx].astype(np.float64)
        trial[np.arange(idx.size), t[idx]] = pn[idx]
        c2 = pb.chi2_fn(trial)
        chi2[idx] = c2
        ibeta = idx // S
        x = -0.5 * pb.beta[ibeta].astype(np.float64) * (c2 - chi2c[idx].astype(np.float64))
        logu = np.log(px.u01(w3[idx]))
        acc = logu <= x
        margin[idx] = logu - x
        accept[idx] = acc
        a = idx[acc]
        pc[a, t[a]] = pn[a]
        chi2c[a] = c2[acc].astype(st.fd)
        np.add.at(st.n_acpt.reshape(P, C), (k[a], a), 1)
    flag[:] = accept.astype(np.int32)
    if burn:
        adjust_steps(pb, st, k)
    if trace is not None:
        trace.append(dict(k=k, inprior=inprior, chi2=chi2, accept=accept,
                          margin=margin, p=pn, sd=sd, lo=pb.pmint[t], hi=pb.pmaxt[t]))


def adjust_steps(pb, st, k):
    """gpu_mcmc.cu:480-526."""
    B, S, P = pb.nbeta, pb.nseq, pb.nprm
    C = B * S
    ci = np.arange(C)
    n_cnt = st.n_cnt.reshape(P, C)
    hist = st.n_hist.reshape(pb.nadj, P, C)
    flag = st.flag.reshape(C)
    pstp = st.pstp.reshape(P, C)
    cnt = n_cnt[k, ci]
    modn = cnt % pb.nadj
    hist[modn, k, ci] = (flag != 0).astype(np.int32)
    sumh = hist[:, k, ci].sum(axis=0)
    rate = np.float32(1.0) * sumh.astype(np.float32) / np.minimum(pb.nadj, cnt).astype(np.float32)
    old = pstp[k, ci].astype(np.float64)
    pstp[k, ci] = np.where(rate < pb.adapt_target, old / 1.01, old * 1.01).astype(st.fd)
    n_cnt[k, ci] = cnt + 1


def exchange(pb, st, it, burn, itr_out=None, trace=None):
    """gpu_mcmc.cu:533-626 (nbeta attempts per sequence) and output recording."""
    B, S, N = pb.nbeta, pb.nseq, pb.nptot
    seq = np.arange(S, dtype=np.uint64)
    ks, accs, margins = [], [], []
    if B > 1:
        for a in range(B):
            w0, w1, _, _ = px.philox4x32(seq, it, a, px.EXCH, *pb.key)
            k = _pick(B - 1, w0,
                      lambda r: px.philox4x32(seq, it, a, px.EXCH + r, *pb.key)[0])
            s = np.arange(S)
            c0 = st.chi2c[k, s].astype(np.float64)
            c1 = st.chi2c[k + 1, s].astype(np.float64)
            db = pb.beta[k + 1].astype(np.float64) - pb.beta[k].astype(np.float64)
            x = -0.5 * db * (c0 - c1)
            logu = np.log(px.u01(w1))
            acc = logu <= x
            sw = s[acc]
            ka = k[acc]
            for arr in (st.tcur, st.chi2c):
                tmp = arr[ka, sw].copy()
                arr[ka, sw] = arr[ka + 1, sw]
                arr[ka + 1, sw] = tmp
            tmp = st.pcur[ka, sw, :].copy()
            st.pcur[ka, sw, :] = st.pcur[ka + 1, sw, :]
            st.pcur[ka + 1, sw, :] = tmp
            if not burn:
                np.add.at(st.n_exch, (ka, sw), 1)
            ks.append(k)
            accs.append(acc)
            margins.append(logu - x)
    if trace is not None:
        trace.append(dict(k=np.array(ks), accept=np.array(accs), margin=np.array(margins)))
    if not burn and itr_out is not None:
        st.pout[:, :, :, itr_out] = np.transpose(st.pcur[:, :, pb.ivar], (2, 0, 1))
        st.tout[:, :, itr_out] = st.tcur
        st.chi2[:, itr_out] = st.chi2c[0, :]


def iteration(pb, st, it, burn, itr_out=None, trace=None):
    for sub in range(pb.nprm):
        mh_substep(pb, st, it, sub, burn, trace=None if trace is None else trace.setdefault("mh", []))
    exchange(pb, st, it, burn, itr_out, trace=None if trace is None else trace.setdefault("ex", []))


def run(pb, nburn, niter, st=None, init=True, trace=None, state_dtype=np.float32):
    """run_mcmc of gpu_mcmc.cu:682-763.  Returns the final State."""
    if st is None:
        st = State(pb, niter, state_dtype)
    if init:
        init_pcur(pb, st)
    for it in range(nburn):
        iteration(pb, st, it, True, trace=trace)
    for itr in range(niter):
        iteration(pb, st, nburn + itr, False, itr_out=itr, trace=trace)
    return st


def geometric_beta(nbeta, beta1=1.0, betan=1e-4):
    """Temperature ladder of src/imgpu.py:412-418 (float32)."""
    beta = np.zeros(nbeta, np.float32)
    bstp = (np.float32(betan) / np.float32(beta1)) ** (1.0 / np.float32(nbeta - 1)) if nbeta > 1 else 1.0
    for i in range(nbeta):
        beta[i] = beta1 * bstp ** i
    return beta
