"""Builds libsnoggpu.so (the C-ABI library with the gfx950 kernels) in-tree.

    python -m snogray_amd.build

All sources go through hipcc for gfx950.  Floating-point flags keep one
IEEE operation per written operation: no FMA contraction, IEEE division
and square root, denormals preserved; packed-FP32 instructions are
disabled.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsnoggpu.so")
SOURCES = ["sgr_kernels.hip", "sgr_api.cpp", "octree_build.cpp", "sgr_jobs.cpp",
           "mlp_kernels.hip", "sgr_guide.cpp", "image_io.cpp", "mesh_io.cpp",
           "sgr_photon.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SGR_OFFLOAD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-fno-gpu-flush-denormals-to-zero",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value"]


def sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False, jobs=8):
    """Compile every source for gfx950 and link libsnoggpu.so."""
    objdir = tempfile.mkdtemp(prefix="sgr_build_")
    objs, procs = [], []
    for s in sources():
        src = os.path.join(CSRC, s)
        obj = os.path.join(objdir, s + ".o")
        cmd = [HIPCC, "-x", "hip", "-c", src, "-o", obj] + FLAGS
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                          stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [l for l in text.splitlines()
                 if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
        if p.returncode != 0:
            failed.append(s)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    os.rmdir(objdir)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
