"""Builds the native parts in-tree.

  libe8q_gpu.so   hipcc, gfx950: HIP kernels + the C ABI (include/e8q_gpu.h)
  _e8qlib*.so     g++ C++17: host E8qIndex over the C ABI + pybind11 module
  oracle          g++: the CPU oracle (test infrastructure) and, when the
                  reference sources are present, the reference library

Every call compiles (no up-to-date shortcut): the caller asked for a build.
"""
import glob
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "e8qlib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_gpu():
    """One object per .hip source, compiled in parallel, then one shared link."""
    from concurrent.futures import ThreadPoolExecutor
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    out = os.path.join(PKG, "libe8q_gpu.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: run([hipcc] + HIP_FLAGS + ["-c", so[0], "-o", so[1]]), zip(srcs, objs)))
    run([hipcc] + HIP_FLAGS + ["-shared"] + objs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_host():
    import pybind11
    srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    if not srcs:
        return
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "_e8qlib" + ext)
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(CSRC, 'host')}"]
    run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fvisibility=hidden",
         "-Wall", "-Wextra"] + inc + srcs + ["-o", out + ".tmp", f"-L{PKG}", "-le8q_gpu",
                                            "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(out + ".tmp", out)


def build_oracle():
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("e8q_oracle.cpp", "me8q_oracle.cpp")]
    out = os.path.join(ROOT, "oracle", "liborc_e8q.so")
    run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp"] + srcs +
        ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    ref = "/root/reference"
    if os.path.isdir(os.path.join(ref, "src")) and shutil.which("make"):
        run(["make", "-s", "-j8", "-f", os.path.join(ROOT, "oracle", "ref.mk"), f"REF={ref}"], cwd=ROOT)


def main(parts=("gpu", "host", "oracle")):
    os.makedirs(PKG, exist_ok=True)
    if "gpu" in parts:
        build_gpu()
    if "host" in parts:
        build_host()
    if "oracle" in parts:
        build_oracle()


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("gpu", "host", "oracle"))
