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


# The tile-pass interpreter branches only on uniform values (the op codes and
# register-bit tests of its program); with the structurizer left off such
# regions, each op updates the register tile in place instead of merging
# copies of it at every join (about a third fewer registers, four fifths
# fewer moves).
TILE_FLAGS = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]

# (source, object, extra flags): the tile-pass kernel once per
# (precision, tile size) so that its variants compile in parallel
ENGINE_UNITS = [("pl_amd.hip", "pl_amd.o", []), ("batch.hip", "batch.o", []), ("mlp.hip", "mlp.o", [])] + [
    ("tiled.hip", f"tiled_{'c128' if d else 'c64'}_{kb}.o",
     [f"-DPL_TILE_DOUBLE={d}", f"-DPL_TILE_KB={kb}", *TILE_FLAGS])
    for d in (1, 0) for kb in ((3, 4, 5) if d else (3, 4, 5, 6))]


def _deps(path, seen=None):
    """The file and every header it includes with quotes, recursively."""
    seen = set() if seen is None else seen
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("#include \""):
                inc = line.split('"')[1]
                _deps(os.path.normpath(os.path.join(os.path.dirname(path), inc)), seen)
    return seen


def build_engine(verbose=False, incremental=False):
    """Each translation unit compiles on its own, all in parallel, then one
    link.  incremental=True skips units whose object is newer than every file
    they include (for development loops; build() always compiles)."""
    out = os.path.join(HERE, "libpl_amd.so")
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    procs, objs = [], []
    for unit, obj, extra in ENGINE_UNITS:
        src = os.path.join(CSRC, unit)
        obj = os.path.join(objdir, obj)
        objs.append(obj)
        if incremental and os.path.exists(obj):
            newest = max(os.path.getmtime(d) for d in _deps(src))
            if os.path.getmtime(obj) > max(newest, os.path.getmtime(__file__)):
                continue
        cmd = [HIPCC, *compile_flags, *extra, "-c", "-o", obj, src]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = None
    for cmd, pr in procs:
        text, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(text)
            failed = failed or cmd
    if failed:
        raise RuntimeError(f"build step failed: {failed[-1]}")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs], verbose)
    return out


def build_stream_probe(verbose=False):
    """libpl_stream.so: the device streaming probe the benchmark's stream mode
    loads (not part of the engine)."""
    out = os.path.join(HERE, "libpl_stream.so")
    _run([HIPCC, *HIP_FLAGS, "-o", out, os.path.join(CSRC, "stream_probe.hip")], verbose)
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
    build_stream_probe(verbose)
    build_bindings(verbose)
    build_oracle(verbose)
