"""Build ``libqsim_mi355x.so`` in place with hipcc for gfx950.

Run ``python -m qibo_amd._build`` (or ``__graft_entry__.build()``). Every HIP
source is compiled to an object with ``--offload-arch=gfx950`` and linked into
one shared library next to this file, where the ctypes loader looks for it.
``measure.hip`` is compiled with ``-ffp-contract=off`` because its arithmetic
must round exactly like numpy's. ``blocked.hip`` keeps uniform branches
unstructurized (its instruction loop switches on a uniform opcode; structurizing
that switch keeps every case's inputs live through the others and costs a copy
of the lane's amplitude registers per instruction).
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_objs")
LIB = os.path.join(HERE, "libqsim_mi355x.so")
ARCH = os.environ.get("QSIM_OFFLOAD_ARCH", "gfx950")

SOURCES = {
    "gates.hip": [],
    "fused.hip": [],
    "measure.hip": ["-ffp-contract=off"],
    "expval.hip": [],
    "blocked.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions=true"],
    "density.hip": ["-ffp-contract=off"],
    "planner.hip": [],
    "compile.hip": ["-ffp-contract=off"],
    "fuse.hip": ["-ffp-contract=off"],
    "compare.hip": [],
    # run-time specialised blocked passes: the device headers' text is
    # embedded (jit_sources.inc, written below) and compiled by hiprtc
    "jit.hip": ["-I" + os.path.join(HERE, "_objs")],
}

# headers compiled again at run time by hiprtc (jit.hip), embedded as strings
JIT_HEADERS = (("kJitQsimDeviceH", "qsim_device.h"), ("kJitBlockedDevH", "blocked_dev.h"))


def _write_jit_sources():
    parts = []
    for name, fname in JIT_HEADERS:
        text = open(os.path.join(CSRC, fname)).read()
        if ")QSIMJIT" in text:
            raise RuntimeError(f"{fname} contains the raw-string delimiter")
        parts.append(f'static const char {name}[] = R"QSIMJIT({text})QSIMJIT";\n')
    out = os.path.join(BUILD, "jit_sources.inc")
    data = "".join(parts)
    if not os.path.exists(out) or open(out).read() != data:
        with open(out, "w") as f:
            f.write(data)

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the MI355X library cannot be built")


def _compile(src, extra):
    hipcc = _hipcc()
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [hipcc, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj, res.stderr


def build(verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    _write_jit_sources()
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        results = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    objs = [o for o, _ in results]
    if verbose:
        for (_, err), src in zip(results, SOURCES):
            if err.strip():
                print(f"[{src}] {err}", file=sys.stderr)
    tmp = LIB + ".tmp"
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(_hipcc()))), "lib")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
           f"-L{rocm_lib}", "-lhiprtc", f"-Wl,-rpath,{rocm_lib}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
