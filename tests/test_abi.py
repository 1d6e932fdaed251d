"""C ABI: every built library exports every function include/mhre.h declares;
host-only entry points and argument validation work without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cuda_mcmc_amd import _lib, build
from cuda_mcmc_amd.engine import Fit
from cuda_mcmc_amd import synthdata

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhre.h")
LIBS = [out for _, out, _ in build.IN_TREE]


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for n in ("mcmc_cuda", "calc_modchi2", "calc_model", "reset_gpu", "list_gpus",
              "mhre_session_create", "mhre_session_run", "mhre_session_download", "mhre_run_batch"):
        assert n in names


@pytest.mark.parametrize("path", LIBS)
def test_library_exports_every_declared_symbol(path):
    assert os.path.isfile(path), "library not built: %s" % path
    L = C.CDLL(path)
    for n in declared_functions():
        assert hasattr(L, n), "%s missing from %s" % (n, path)


def test_host_entry_points():
    L = _lib.load(LIBS[0])
    assert L.mhre_version() >= 10000
    assert L.mhre_native_eht() == 1
    assert L.mhre_user_model().decode().endswith("model.cuh")
    for other in LIBS[1:]:
        Lo = _lib.load(other)
        assert Lo.mhre_native_eht() == 0
    # the product's Philox against the Random123 known-answer vectors
    from tests.test_oracle_philox import KAT
    L.mhre_philox4x32_10.restype = None
    for ctr, key, want in KAT:
        c = (C.c_uint * 4)(*ctr)
        k = (C.c_uint * 2)(*key)
        o = (C.c_uint * 4)()
        L.mhre_philox4x32_10(c, k, o)
        assert tuple(o) == want


def _fit():
    d = synthdata.ring_gauss_dataset(nvis=100, ncph=30, nsta=5, seed=3)
    idat = np.array([d["nvis"], 0, 0], np.int32)
    return Fit(synthdata.EHT_PDESCR, synthdata.EHT_PMIN, synthdata.EHT_PMAX, d["coor"], d["dat"], d["std2r"],
               nbeta=2, nseq=2, nburn=1, niter=1, model="sgra9", imodel=1, icoor=d["icoor"], idat=idat, npass=2)


def _create(fits):
    L = _lib.load(LIBS[0])
    specs = (_lib.MhreFitSpec * len(fits))(*[f.spec() for f in fits])
    h = C.c_void_p()
    rc = L.mhre_session_create(specs, len(fits), None, C.byref(h))
    return rc, L.mhre_last_error().decode()


def test_validation_errors_are_codes_not_exits():
    L = _lib.load(LIBS[0])
    h = C.c_void_p()
    assert L.mhre_session_create(None, 0, None, C.byref(h)) == -1
    f = _fit()
    f.invar[3] = 7                       # inconsistent ivar / invar
    rc, msg = _create([f])
    assert rc == -1 and "invar" in msg
    f = _fit()
    f.n_cnt[0, 0, 0] = -1                # would index n_hist out of bounds
    rc, msg = _create([f])
    assert rc == -1 and "n_cnt" in msg
    f = _fit()
    f.icoor[5] = 10 ** 6                 # closure index outside the visibilities
    rc, msg = _create([f])
    assert rc == -1 and "icoor" in msg
    f = _fit()
    f.idat[0] = 0
    rc, msg = _create([f])
    assert rc == -1 and "idat" in msg


def _prototypes():
    """name -> number of parameters, from the prototypes of include/mhre.h."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?(?:int|void|char)\s*\*?\s*(\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else len(args.split(","))
    return out


def test_ctypes_declarations_match_the_header():
    """Every argtypes list the Python layer declares has the header's arity."""
    L = _lib.load(LIBS[0])
    protos = _prototypes()
    checked = 0
    for name, n in protos.items():
        f = getattr(L, name)
        if f.argtypes is None:
            continue
        assert len(f.argtypes) == n, "%s: ctypes declares %d arguments, header %d" % (name, len(f.argtypes), n)
        checked += 1
    assert checked >= 15
    assert protos["mcmc_cuda"] == 44
