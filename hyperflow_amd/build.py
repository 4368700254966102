"""Builds the gfx950 library hyperflow_amd/libhyperflow_amd.so with hipcc.

Kernels are compiled per arithmetic type, both with -ffp-contract=off: in
FP64 that is the reference's rounding; in FP32 it keeps results independent
of the tile split (with contraction the compiler fuses the warm-up and the
loop-body copies of one reconstruction differently).  Packed-FP32
instructions are disabled for all device code.  The host C ABI is compiled with -ffp-contract=off as well, since
it evaluates FP64 timestep arithmetic on the CPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libhyperflow_amd.so")
ARCH = os.environ.get("HF_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-std=c++17", "-fPIC", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
# Atomics whose address is wave-uniform are combined within the wave by the
# compiler; the DPP strategy does it with a few cross-lane adds instead of a
# scalar loop over the 64 lanes (the work-queue ticket of every item).
DEVICE = [f"--offload-arch={ARCH}", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-atomic-optimizer-strategy=DPP"]

# (source, flags, hidden visibility): the C++ drop-in classes are exported
UNITS = [
    # The FP64 row loop is VALU-bound with one or two waves per SIMD: the
    # scheduler's max-ILP strategy (up to 253 VGPRs, still two waves) runs it
    # 1-2 % faster; the FP32 loop, at three waves per SIMD, is 4 % slower
    # with it and keeps the default.
    ("hf_kernels_f64.hip", DEVICE + ["-ffp-contract=off", "-mllvm", "-amdgpu-sched-strategy=max-ilp"], True),
    ("hf_kernels_f32.hip", DEVICE + ["-ffp-contract=off"], True),
    ("hf_cnn.hip", DEVICE + ["-ffp-contract=off"], True),
    ("hf_noalloc.hip", DEVICE, True),
    ("hf_capi.cpp", ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-O2", "-ffp-contract=off"], True),
    ("hf_learned.cpp", ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-O2"], True),
    ("hf_uns.hip", DEVICE + ["-ffp-contract=off"], True),
    ("hf_uns.cpp", ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-O2", "-ffp-contract=off"], True),
    ("host/hf_json.cpp", ["-x", "c++", "-O2"], False),
    ("host/hf_host.cpp", ["-x", "c++", "-O2", "-ffp-contract=off"], False),
]
CLI = os.path.join(HERE, "bin", "hyperflow")
CLIS = ((CLI, "host/hf_cli.cpp"), (os.path.join(HERE, "bin", "hyperflow_uns"), "host/hf_uns_cli.cpp"))
# test driver of the reference-typed C++ surface (tests/cpp), linked like the CLIs
TEST_DRIVERS = ((os.path.join(ROOT, "tests", "bin", "ref_loop"), os.path.join(ROOT, "tests", "cpp", "ref_loop.cpp")),)


def _compile(unit, flags, hidden, outdir):
    src = unit if os.path.isabs(unit) else os.path.join(CSRC, unit)
    obj = os.path.join(outdir, os.path.basename(unit) + ".o")
    cmd = [HIPCC] + COMMON + (["-fvisibility=hidden"] if hidden else []) + flags + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    msgs = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n" + "\n".join(msgs))
    return obj


def build(verbose: bool = False) -> str:
    outdir = os.path.join(HERE, "_build")
    os.makedirs(outdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        objs = list(ex.map(lambda u: _compile(u[0], u[1], u[2], outdir), UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    # the `hyperflow` and `hyperflow_uns` command-line drivers, linked against the library
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    for exe, src in CLIS:
        cli_obj = _compile(src, ["-x", "c++", "-O2"], False, outdir)
        cmd = [HIPCC, "-o", exe + ".tmp", cli_obj, f"-L{HERE}", "-lhyperflow_amd", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{src}: link failed:\n" + r.stdout)
        os.replace(exe + ".tmp", exe)
    for exe, src in TEST_DRIVERS:
        os.makedirs(os.path.dirname(exe), exist_ok=True)
        obj = _compile(src, ["-x", "c++", "-O2"], False, outdir)
        cmd = [HIPCC, "-o", exe + ".tmp", obj, f"-L{HERE}", "-lhyperflow_amd",
               "-Wl,-rpath,$ORIGIN/../../hyperflow_amd"]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{src}: link failed:\n" + r.stdout)
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {LIB} and {', '.join(e for e, _ in CLIS)}")
    return LIB


def build_oracle(verbose: bool = False) -> None:
    """Compile oracle/ (and oracle/_ref when the reference checkout exists)."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    if verbose and r.stdout.strip():
        print(r.stdout)


if __name__ == "__main__":
    build(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
