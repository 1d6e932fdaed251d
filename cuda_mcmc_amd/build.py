"""Build the engine's shared libraries with hipcc for gfx950.

The reference rebuilds its static library and Cython module per project with
that project's model.cuh (``src/Makefile:37-45``, ``template/Makefile:12-15``).
Here one shared library is built per model directory:

    python -m cuda_mcmc_amd.build                     # every in-tree library
    python -m cuda_mcmc_amd.build --model-dir DIR     # DIR/libmhre.so from DIR/model.cuh

The package default library (``cuda_mcmc_amd/lib/libmhre.so``) is built from
``examples/sgra_model_fit/model.cuh`` with the native fused ring+Gaussian
kernels enabled for imodel 1 and 2.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
EXAMPLES = os.path.join(ROOT, "examples")
ARCH = os.environ.get("MHRE_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from forming packed fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + CSRC, "-I" + INCLUDE,
]

# (model dir, output library, native ring+Gaussian kernels for imodel 1/2)
IN_TREE = [
    (os.path.join(EXAMPLES, "sgra_model_fit"), os.path.join(LIBDIR, "libmhre.so"), True),
    (os.path.join(EXAMPLES, "search_Rosenbrock"), os.path.join(EXAMPLES, "search_Rosenbrock", "libmhre.so"), False),
    (os.path.join(EXAMPLES, "solvsys"), os.path.join(EXAMPLES, "solvsys", "libmhre.so"), False),
    (os.path.join(EXAMPLES, "template"), os.path.join(EXAMPLES, "template", "libmhre.so"), False),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), out))
    return out


def _obj(name, src, extra):
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, name)
    _run([HIPCC] + FLAGS + extra + ["-c", src, "-o", out])
    return out


def build_library(model_dir, out, native_eht=False, tag=None):
    """Compile engine + generic(model.cuh of model_dir) + api and link ``out``."""
    model_dir = os.path.abspath(model_dir)
    if not os.path.isfile(os.path.join(model_dir, "model.cuh")):
        raise FileNotFoundError(os.path.join(model_dir, "model.cuh"))
    tag = tag or os.path.basename(model_dir.rstrip("/"))
    name = os.path.basename(model_dir.rstrip("/")) + "/model.cuh"
    jobs = [
        ("engine.o", os.path.join(CSRC, "engine.hip"), []),
        ("grad.o", os.path.join(CSRC, "grad.hip"), []),
        ("generic_%s.o" % tag, os.path.join(CSRC, "generic.hip"),
         ["-I" + model_dir, '-DMHRE_MODEL_NAME="%s"' % name]),
        ("api_%s.o" % ("native" if native_eht else "plain"), os.path.join(CSRC, "api.hip"),
         ["-DMHRE_NATIVE_EHT=1"] if native_eht else []),
    ]
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda j: _obj(*j), jobs))
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_program(src, out, verbose=True):
    """A stand-alone HIP program (test and measurement helpers) with the library's flags."""
    if verbose:
        print("building", os.path.relpath(out, ROOT), flush=True)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    _run([HIPCC] + FLAGS + [src, "-o", out])
    return out


def build_all(verbose=True):
    built = []
    for model_dir, out, native in IN_TREE:
        if verbose:
            print("building", os.path.relpath(out, ROOT), flush=True)
        built.append(build_library(model_dir, out, native))
    return built


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--model-dir", help="project directory holding model.cuh")
    ap.add_argument("--out", help="output library (default MODEL_DIR/libmhre.so)")
    ap.add_argument("--native-eht", action="store_true",
                    help="route imodel 1/2 to the native fused ring+Gaussian kernels")
    a = ap.parse_args(argv)
    if a.model_dir:
        out = a.out or os.path.join(a.model_dir, "libmhre.so")
        print(build_library(a.model_dir, out, a.native_eht))
    else:
        for p in build_all():
            print(p)


if __name__ == "__main__":
    sys.exit(main())
