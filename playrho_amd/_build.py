"""Builds the in-tree native libraries.

* playrho_amd/libplayrho_amd.so: the C ABI (host C++17) and the gfx950 HIP
  kernels, compiled with hipcc for gfx950 only.  Float semantics: no FMA
  contraction, IEEE denormals, correctly rounded division and square root.
* playrho_amd/libplayrho_amd_torch.so: the libtorch operator
  torch.ops.playrho_amd.billiards_loss (a torch::autograd::Function over the
  adjoint's C ABI), compiled with g++ against the installed libtorch.
* oracle/_build/liboracle.so: the CPU restatement used as the checker.
* oracle/_ref/: the reference float build and golden driver (only where the
  reference checkout is present).
"""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libplayrho_amd.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout[-4000:], r.stderr[-8000:]))
    return r


def build_library(out=LIB):
    objs = []
    tmp = os.path.join(PKG, "_build")
    os.makedirs(tmp, exist_ok=True)
    # The translation units compile in parallel.
    procs = []
    for src in ("step_kernel.hip", "big_world.hip", "adjoint.hip", "host_world.cpp"):
        obj = os.path.join(tmp, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        so, se = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), so[-4000:], se[-8000:]))
    tmp_out = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_out] + objs)
    os.replace(tmp_out, out)
    return out


TORCH_LIB = os.path.join(PKG, "libplayrho_amd_torch.so")


def build_torch_op(out=TORCH_LIB):
    import torch
    import torch.utils.cpp_extension as ce
    abi = "1" if torch._C._GLIBCXX_USE_CXX11_ABI else "0"
    inc = []
    for p in ce.include_paths():
        inc += ["-isystem", p]
    libs = []
    for p in ce.library_paths():
        libs += ["-L" + p, "-Wl,-rpath," + p]
    tmp_out = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-D_GLIBCXX_USE_CXX11_ABI=" + abi,
          "-I" + os.path.join(ROOT, "include")] + inc +
         [os.path.join(CSRC, "torch_adjoint.cpp"), "-o", tmp_out] + libs +
         ["-ltorch", "-ltorch_cpu", "-lc10", "-L" + PKG, "-lplayrho_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp_out, out)
    return out


def build_oracle(out=ORACLE_LIB):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          os.path.join(ROOT, "oracle", "cpu_ref.cpp"), "-o", out, "-lpthread"])
    return out


def build_reference():
    """Builds oracle/_ref from /root/reference when present (here only)."""
    if not os.path.isdir("/root/reference/Library/source"):
        return None
    jobs = str(min(16, os.cpu_count() or 4))
    _run(["make", "-f", "oracle/Makefile.ref", "-j" + jobs, "all"], cwd=ROOT)
    return os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def build_all():
    build_library()
    build_torch_op()
    build_oracle()
    build_reference()
