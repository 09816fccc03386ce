"""Build recipe for the native parts, run in-tree.

  graphindor_amd/lib/libgm_hip.so    HIP kernels + C ABI (hipcc, gfx950)
  graphindor_amd/lib/libgraphset.so  host C++ library (Graph / Pattern / Schedule_IEP ...)
  graphindor_amd/lib/gpu_graph, gpu_clique, gpu_mc   command lines
  oracle/liboracle.so                CPU oracle (test infrastructure)
  oracle/_ref/                       the reference CPU path, when its sources are present

Everything is compiled on every call (the build check compiles even when
libraries exist).  Usage: python -m graphindor_amd.build [-v]
"""
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "graphindor_amd")
LIB = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
HOST_INC = os.path.join(CSRC, "host", "include")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GM_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 instruction selection is disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-sign-compare"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")
    elif verbose and r.stdout.strip():
        print(r.stdout)


def build_device(verbose=False):
    """Compiles the device objects in parallel -- gm_engine.hip, gm_mlp.hip and
    one object per enumerate_kernel<MODE, NMAX, WPE> instantiation of
    enum_instance.hip, the deepest first since they take longest -- and links
    libgm_hip.so."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB, exist_ok=True)
    obj_dir = os.path.join(LIB, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    dev = os.path.join(CSRC, "device")
    jobs = []
    for src in sorted(glob.glob(os.path.join(dev, "*.hip"))):
        base = os.path.splitext(os.path.basename(src))[0]
        if base == "enum_instance":
            for mode in (0, 1):
                # the pattern rule in 2- and 4-wave-per-SIMD builds, the clique rule in the 2-wave one
                for wpe in ((2, 4) if mode == 0 else (2,)):
                    for nmax in (4, 5, 6, 7):
                        jobs.append((src, os.path.join(obj_dir, f"{base}_{mode}_{nmax}_{wpe}.o"),
                                     [f"-DGMK_MODE={mode}", f"-DGMK_NMAX={nmax}", f"-DGMK_WPE={wpe}"]))
        else:
            jobs.append((src, os.path.join(obj_dir, base + ".o"), []))
    jobs.sort(key=lambda j: -sum(int(d.split("=")[1]) * (10 if "NMAX" in d else 0) + (5 if d == "-DGMK_MODE=0" else 0)
                                 for d in j[2]))

    def compile_one(job):
        src, obj, defs = job
        _run([HIPCC, *HIP_FLAGS, *defs, f"-I{INC}", f"-I{dev}", "-c", src, "-o", obj], verbose)
        return obj
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(compile_one, jobs))
    out = os.path.join(LIB, "libgm_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_host(verbose=False):
    out = os.path.join(LIB, "libgraphset.so")
    srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    _run(["g++", *HOST_FLAGS, "-shared", f"-I{INC}", f"-I{HOST_INC}", *srcs,
          f"-L{LIB}", "-lgm_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    for cli in sorted(glob.glob(os.path.join(CSRC, "cli", "*.cpp"))):
        exe = os.path.join(LIB, os.path.splitext(os.path.basename(cli))[0])
        _run(["g++", *HOST_FLAGS, f"-I{INC}", f"-I{HOST_INC}", cli, f"-L{LIB}", "-lgraphset", "-lgm_hip",
              "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"], verbose)
        os.replace(exe + ".tmp", exe)
    return out


def build_oracle(verbose=False):
    out = os.path.join(ORACLE, "liboracle.so")
    _run(["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall",
          os.path.join(ORACLE, "gs_oracle.c"), "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    ref = os.environ.get("GRAPHSET_REF", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src")):
        _run(["sh", os.path.join(ORACLE, "build_ref.sh")], verbose)
    return out


def build_all(verbose=False):
    build_device(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("build ok")
