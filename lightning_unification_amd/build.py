"""In-tree build of the native parts (no CMake, no JIT cache).

  libpl_amd.so             HIP engine + C-ABI (hipcc, gfx950)
  lightning_qubit_ops*.so  pybind11 mirror of Lightning-Qubit's module (g++),
                           linked against libpl_amd.so
  oracle/liblq_oracle.so   CPU oracle restatement (test infrastructure)
  oracle/_ref/libref_lq.so reference LQ sources compiled in place, only where
                           the reference tree exists (test infrastructure)
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PL_AMD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared", "-Wall",
    # packed FP32 (v_pk_*_f32) stays off: hipcc forms it from plain float code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return r


def build_engine(verbose=False):
    out = os.path.join(HERE, "libpl_amd.so")
    _run([HIPCC, *HIP_FLAGS, "-o", out, os.path.join(CSRC, "pl_amd.hip")], verbose)
    return out


def build_bindings(verbose=False):
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "lightning_qubit_ops" + suffix)
    src = os.path.join(CSRC, "host", "bindings.cpp")
    if not os.path.exists(src):
        return None
    cmd = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-Wall",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(CSRC, 'host')}",
           src, "-o", out, f"-L{HERE}", "-lpl_amd", "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    return out


def build_oracle(verbose=False):
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-B", "-s", "-C", odir, "-j8"], verbose)
    if os.path.isdir(REFERENCE):
        _run(["make", "-B", "-s", "-C", odir, "-j8", "ref", f"REF={REFERENCE}"], verbose)


def build_all(verbose=False):
    build_engine(verbose)
    build_bindings(verbose)
    build_oracle(verbose)
