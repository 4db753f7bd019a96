"""In-tree build of the gfx950 libraries.

  lib/libtsp_hip.so   hipcc --offload-arch=gfx950: csrc/*.hip (kernels + C ABI)
  lib/libtsp_host.so  g++: host/*.cpp (scene builder, SimBackend_HIP), linked to libtsp_hip.so
  lib/sim_backend_roundtrip  g++: the C++ driver of SimBackend_HIP that the tests run
  lib/math_selftest  hipcc: GPU check of the short exact sqrt / quotient against the general lowering

Floating point: -ffp-contract=off everywhere (the reference's ISO C++ build contracts nothing;
every fused multiply-add is an explicit fmaf), correctly rounded f32 division and square root,
and no packed-FP32 instructions.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["kernels_search.hip", "kernels_sph.hip", "kernels_stats.hip", "kernels_contact.hip", "kernels_export.hip", "kernels_build.hip", "tsp_capi.hip"]
HOST_SOURCES = ["scene_plan.cpp", "scene_builder.cpp", "sim_backend_hip.cpp", "state_io.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    # calls inside the library bind to its own definitions (another build loaded next to it
    # cannot interpose them)
    "-Wl,-Bsymbolic",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def hip_command(out, defines=()):
    """The hipcc command that builds libtsp_hip.so at `out` (with extra -D defines)."""
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    # the scene plan is shared with the host builder (plain C++, compiled here too)
    srcs = [os.path.join(HERE, "csrc", s) for s in HIP_SOURCES] + [os.path.join(HERE, "host", "scene_plan.cpp")]
    return [hipcc] + HIPCC_FLAGS + list(defines) + ["-I" + INC, "-I" + os.path.join(HERE, "csrc"),
                                                     "-I" + os.path.join(HERE, "host"), "-shared", "-o", out] + srcs


def build_hip(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libtsp_hip.so")
    r = _run(hip_command(out + ".tmp"))
    os.replace(out + ".tmp", out)
    if verbose and r.stderr:
        print(r.stderr)
    return out


def build_host(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libtsp_host.so")
    srcs = [os.path.join(HERE, "host", s) for s in HOST_SOURCES if os.path.exists(os.path.join(HERE, "host", s))]
    cmd = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
           "-I" + INC, "-I" + os.path.join(HERE, "host"), "-o", out + ".tmp"] + srcs + \
          ["-L" + LIB, "-ltsp_hip", "-Wl,-rpath,$ORIGIN"]
    r = _run(cmd)
    os.replace(out + ".tmp", out)
    if verbose and r.stderr:
        print(r.stderr)
    return out


def build_driver(verbose=False):
    out = os.path.join(LIB, "sim_backend_roundtrip")
    src = os.path.join(ROOT, "tests", "cpp", "sim_backend_roundtrip.cpp")
    cmd = ["g++", "-O2", "-std=c++20", "-Wall", "-I" + INC, "-o", out + ".tmp", src,
           "-L" + LIB, "-ltsp_host", "-ltsp_hip", "-Wl,-rpath,$ORIGIN"]
    r = _run(cmd)
    os.replace(out + ".tmp", out)
    if verbose and r.stderr:
        print(r.stderr)
    return out


def build_math_selftest(verbose=False):
    """lib/math_selftest: the GPU check of the short exact sqrt / quotient against the compiler's"""
    out = os.path.join(LIB, "math_selftest")
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    src = os.path.join(ROOT, "tests", "cpp", "math_selftest.hip")
    cmd = [hipcc] + HIPCC_FLAGS + ["-I" + INC, "-I" + os.path.join(HERE, "csrc"), "-o", out + ".tmp", src]
    r = _run(cmd)
    os.replace(out + ".tmp", out)
    if verbose and r.stderr:
        print(r.stderr)
    return out


def build_all(verbose=False):
    build_hip(verbose)
    build_host(verbose)
    build_driver(verbose)
    build_math_selftest(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
