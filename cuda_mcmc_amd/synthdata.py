"""Synthetic workloads for the MH-RE engine.

ring_gauss_dataset()  an EHT-like observation of the 9-parameter ring+Gaussian
                      source: stations at random sites, Earth-rotation uv
                      tracks at 230 GHz, per-time closure triangles built
                      with the reference's closure combination
                      phi(a1a2) + phi(a1a3) - phi(a2a3)
                      (``src/imgpu.py:1294-1306``), amplitude and closure-phase
                      noise, packed in the reference's ``coor/dat/std2r/icoor/
                      idat`` layout (``src/imgpu.py:759-831``).
quadsys_dataset()     random systems of quadratic equations with a planted
                      root (the large form of ``examples/solvsys``).
EHT_PRIOR             the 9-parameter prior box of the benchmark fits.

Noise: amplitude sigma = sqrt((f |V|)^2 + floor^2) with f ~ U[1%, 5%] per
visibility and a thermal floor (default 1e-3 Jy) so that no datum has a
vanishing sigma; closure-phase sigma ~ U[2, 10] degrees per triangle.
"""
import numpy as np

C_LIGHT = 299792458.0
EARTH_R = 6.371e6
UAS2NRAD = np.pi / 180.0 / 3600.0 * 1e3

#                 Zsp   Re    rq   ecc  fade  gax         aq   gq   th
EHT_PMIN = np.array([0.5, 20., 0.3, 0.0, 0.0, 50. / 30., 0.5, 0.0, -np.pi], np.float32)
EHT_PMAX = np.array([3.0, 30., 0.9, 0.5, 1.0, 150. / 20., 1.0, 0.5, np.pi], np.float32)
EHT_PDESCR = np.array([1, 1, 1, 1, 1, 1, 1, 1, 2], np.int32)


def draw_truth(rng):
    """Truth parameters: uniform in the stated ranges; gax = FWHM/Re."""
    Zsp = rng.uniform(0.5, 3.0)
    Re = rng.uniform(20.0, 30.0)
    rq = rng.uniform(0.3, 0.9)
    ecc = rng.uniform(0.0, 0.5)
    fade = rng.uniform(0.0, 1.0)
    fwhm = rng.uniform(50.0, 150.0)
    aq = rng.uniform(0.5, 1.0)
    gq = rng.uniform(0.0, 0.5)
    th = rng.uniform(0.0, 2 * np.pi)
    th = th - 2 * np.pi * np.rint(th / (2 * np.pi))
    return np.array([Zsp, Re, rq, ecc, fade, fwhm / Re, aq, gq, th], np.float64)


def _station_tracks(rng, nsta, dec, hours, dt_s, max_glam, freq_hz):
    """uv tracks (Glambda) of all station pairs with the source above 10 deg."""
    lam = C_LIGHT / freq_hz
    # sites spread over the hemisphere facing the source; the array is scaled
    # down if its longest baseline would exceed max_glam
    lat = np.radians(rng.uniform(-60.0, 40.0, nsta))
    lon = np.radians(rng.uniform(-70.0, 70.0, nsta))
    xyz = EARTH_R * np.stack([np.cos(lat) * np.cos(lon),
                              np.cos(lat) * np.sin(lon),
                              np.sin(lat)], axis=1)
    dmax = max(np.linalg.norm(xyz[i] - xyz[j]) for i in range(nsta) for j in range(i + 1, nsta))
    xyz *= min(1.0, max_glam * 1e9 * lam / dmax)
    nt = int(round(hours * 3600.0 / dt_s))
    H0 = -0.5 * hours / 24.0 * 2 * np.pi
    H = H0 + np.arange(nt) * dt_s / 86164.0 * 2 * np.pi
    recs = []
    for it in range(nt):
        h = H[it]
        # elevation of the source at each station
        s = np.array([np.cos(dec) * np.cos(h + lon[i]) * np.cos(lat[i]) + np.sin(dec) * np.sin(lat[i])
                      for i in range(nsta)])
        up = np.nonzero(s > np.sin(np.radians(10.0)))[0]
        for a in range(len(up)):
            for b in range(a + 1, len(up)):
                i, j = up[a], up[b]
                L = (xyz[j] - xyz[i]) / lam
                u = np.sin(h) * L[0] + np.cos(h) * L[1]
                v = -np.sin(dec) * np.cos(h) * L[0] + np.sin(dec) * np.sin(h) * L[1] + np.cos(dec) * L[2]
                recs.append((it, i + 1, j + 1, u / 1e9, v / 1e9))
    return recs


def ring_gauss_dataset(nvis=2000, ncph=600, nsta=8, seed=2, truth=None,
                       hours=6.0, dt_s=60.0, max_glam=8.0, freq_hz=230e9,
                       amp_floor=1e-3, true_closure=False):
    """Synthetic 9-parameter ring+Gaussian observation.

    Returns a dict with the truth, per-visibility arrays (u, v, amp, eamp,
    phase, tsec, blin), closures (cphs, ecph, cpix) and the packed reference
    arrays (coor, dat, std2r, icoor, idat) for nbeta*nseq-independent use.
    """
    from .models_fp64 import sgra9_vis
    rng = np.random.default_rng(seed)
    if truth is None:
        truth = draw_truth(rng)
    dec = np.radians(-29.0) + rng.uniform(-0.3, 0.3)
    # finer sampling when the requested size needs more records than the track gives
    per_time = nsta * (nsta - 1) // 2
    while int(hours * 3600 / dt_s) * per_time * 0.5 < nvis:
        dt_s /= 2.0
    recs = _station_tracks(rng, nsta, dec, hours, dt_s, max_glam, freq_hz)
    bytime = {}
    for r in recs:
        bytime.setdefault(r[0], []).append(r)
    times = np.array(sorted(bytime))
    order = rng.permutation(times)
    chosen, count = [], 0
    for t in order:
        if count >= nvis:
            break
        chosen.append(t)
        count += len(bytime[t])
    chosen.sort()
    rows = [r for t in chosen for r in bytime[t]]
    # drop trailing records of the last chosen time to hit nvis exactly
    rows = rows[:nvis]
    nv = len(rows)
    u = np.array([r[3] for r in rows])
    v = np.array([r[4] for r in rows])
    tsec = np.array([r[0] * dt_s for r in rows])
    blin = np.array([(r[1], r[2]) for r in rows], np.int32)
    V = sgra9_vis(u, v, truth)
    f = rng.uniform(0.01, 0.05, nv)
    eamp = np.sqrt((f * np.abs(V)) ** 2 + amp_floor ** 2)
    amp = np.abs(V) + rng.normal(0.0, 1.0, nv) * eamp
    phase = np.angle(V)
    # closure triangles at every time: stations a1<a2<a3, baselines (a1a2),(a1a3),(a2a3)
    tri = []
    start = 0
    while start < nv:
        end = start
        while end < nv and tsec[end] == tsec[start]:
            end += 1
        idx = {(blin[i, 0], blin[i, 1]): i for i in range(start, end)}
        ants = sorted(set(blin[start:end].ravel()))
        for x in range(len(ants)):
            for y in range(x + 1, len(ants)):
                for z in range(y + 1, len(ants)):
                    a1, a2, a3 = ants[x], ants[y], ants[z]
                    if (a1, a2) in idx and (a1, a3) in idx and (a2, a3) in idx:
                        tri.append((idx[(a1, a2)], idx[(a1, a3)], idx[(a2, a3)]))
        start = end
    tri = np.array(tri, np.int32).reshape(-1, 3)
    if ncph > len(tri):
        raise ValueError("only %d closure triangles available, %d requested" % (len(tri), ncph))
    sel = np.sort(rng.choice(len(tri), ncph, replace=False))
    cpix = tri[sel]
    if true_closure:   # opt-in gain-invariant closure phi12 + phi23 - phi13
        cpix = cpix[:, [0, 2, 1]]
    ecph = np.radians(rng.uniform(2.0, 10.0, ncph))
    cph_true = phase[cpix[:, 0]] + phase[cpix[:, 1]] - phase[cpix[:, 2]]
    cph = cph_true + rng.normal(0.0, 1.0, ncph) * ecph
    cph = np.arctan2(np.sin(cph), np.cos(cph))
    return pack_eht(u, v, amp, eamp, cph, ecph, cpix, truth=truth, phase=phase,
                    tsec=tsec, blin=blin)


def pack_eht(u, v, amp, eamp, cphs, ecph, cpix, **extra):
    """Pack per-datum arrays into the reference's coor/dat/std2r/icoor/idat.

    Layout of src/imgpu.py:759-831 with cpExt = 0: coor = [u0,v0,u1,v1,...],
    dat = [amp..., cphs...], std2r = [1/eamp^2..., 1/ecph^2...],
    icoor = cpix.flatten(), idat = [nvis, nchi2m, cpExt] (nchi2m is filled in
    by the caller that knows nbeta*nseq; 0 here).
    """
    nvis = len(amp)
    ncph = len(cphs)
    d = dict(extra)
    d.update(
        u=np.asarray(u, np.float32), v=np.asarray(v, np.float32),
        amp=np.asarray(amp, np.float32), eamp=np.asarray(eamp, np.float32),
        cphs=np.asarray(cphs, np.float32), ecph=np.asarray(ecph, np.float32),
        cpix=np.asarray(cpix, np.int32).reshape(-1, 3), nvis=nvis, ncph=ncph)
    coor = np.empty(2 * nvis, np.float32)
    coor[0::2] = d["u"]
    coor[1::2] = d["v"]
    dat = np.concatenate([d["amp"], d["cphs"]]).astype(np.float32)
    std2r = np.concatenate([1.0 / d["eamp"].astype(np.float64) ** 2,
                            1.0 / d["ecph"].astype(np.float64) ** 2]).astype(np.float32)
    d.update(coor=coor, dat=dat, std2r=std2r,
             icoor=d["cpix"].ravel().astype(np.int32),
             idat=np.array([nvis, 0, 0], np.int32))
    return d


def quadsys_dataset(nsys=1024, neq=8, n=8, seed=5, root_box=2.0):
    """Random quadratic systems x^T Q_e x + b_e . x + c_e = 0 with a planted root.

    Q, b ~ N(0,1); x* ~ U[-root_box, root_box]^n; c_e chosen so F(x*) = 0.
    Returns Q [nsys,neq,n,n], b [nsys,neq,n], c [nsys,neq], xstar [nsys,n] (float64).
    """
    rng = np.random.default_rng(seed)
    Q = rng.normal(size=(nsys, neq, n, n))
    b = rng.normal(size=(nsys, neq, n))
    xs = rng.uniform(-root_box, root_box, size=(nsys, n))
    c = -(np.einsum("si,seij,sj->se", xs, Q, xs) + np.einsum("si,sei->se", xs, b))
    return dict(Q=Q, b=b, c=c, xstar=xs)
