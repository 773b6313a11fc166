"""Build the engine in-tree: cds_amd/libcds_amd.so and cds_amd/bin/cds_main.

Every HIP source is compiled for gfx950 with hipcc; host C++ (which only
calls the HIP runtime API) is compiled with g++.  All objects use
-ffp-contract=off so the engine's float bound/search arithmetic rounds as
written.
"""
import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcds_amd.so")
BIN = os.path.join(HERE, "bin", "cds_main")

HIP_SOURCES = ["scan.hip", "dag.hip", "list.hip", "peel.hip", "glue.hip", "flow.hip",
               "flow_grid.hip", "batch.hip", "noalloc.hip", "prune.hip"]
HOST_SOURCES = ["graph_io.cpp", "engine.cpp", "server.cpp", "report.cpp", "tune.cpp"]
ARCH = os.environ.get("CDS_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-result"]
DEVICE_ONLY = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_ONLY = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]


def _hipcc():
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.isabs(cand) and os.path.exists(cand):
            return cand
    return "hipcc"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


# Kernels launched cooperatively (all workgroups co-resident, spinning on
# grid barriers) must not spill to scratch memory: the build fails if the
# compiler reports any scratch for them.
COOPERATIVE_KERNELS = ("k_peel_persistent", "k_grid_flow", "k_flow_batch",
                       "k_replay_grid")


def _check_scratch(src, stderr):
    name = None
    for line in stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name and int(m.group(1)) > 0 and \
                any(k in name for k in COOPERATIVE_KERNELS):
            raise RuntimeError(f"{src}: cooperative kernel {name} uses "
                               f"{m.group(1)} bytes/lane of scratch")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([hipcc, f"--offload-arch={ARCH}", *COMMON, *DEVICE_ONLY,
                     "-Rpass-analysis=kernel-resource-usage",
                     "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SOURCES + ["main.cpp"]:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++", *COMMON, *HOST_ONLY, "-c", os.path.join(CSRC, src),
                     "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for cmd, r in zip(jobs, ex.map(_run, jobs)):
            if "-Rpass-analysis=kernel-resource-usage" in cmd:
                _check_scratch(cmd[-3], r.stderr)
            elif verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    objs = [os.path.join(BUILD, s + ".o") for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lpthread"])
    os.replace(tmp, LIB)
    _run([hipcc, "-o", BIN, os.path.join(BUILD, "main.cpp.o"), f"-L{HERE}",
          "-lcds_amd", "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
