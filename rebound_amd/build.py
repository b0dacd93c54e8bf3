"""Builds librebound_amd.so (HIP kernels + host runtime) for gfx950, in-tree.

    python -m rebound_amd.build            # library only
    python -m rebound_amd.build --all      # library + the CPU checker (oracle/)

Every translation unit is compiled by hipcc with:
  -ffp-contract=off   the reference is contraction-free; fused multiply-adds
                      appear only where a kernel writes fma() explicitly
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated
  -fvisibility=hidden only the REB_API (reb_*) entry points are exported
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librebound_amd.so")
ARCH = os.environ.get("REBOUND_AMD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-fvisibility=hidden", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def compile_one(src):
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    # .cpp files hold host code only (they include the HIP runtime header for
    # its API): compiled as the host side of HIP, no device object.
    cmd = [hipcc()] + COMMON + ["-x", "hip"] + DEVICE
    if src.endswith(".cpp"):
        cmd.append("--offload-host-only")
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), res.stdout, res.stderr))
    # the host pass does not know the packed-fp32 feature: drop that warning
    noise = [ln for ln in res.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
             and "not a recognized feature" not in ln]
    return obj, "\n".join(noise)


def build_library(verbose=False):
    srcs = sources()
    jobs = min(8, len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, warn in ex.map(compile_one, srcs):
            objs.append(obj)
            if warn and verbose:
                print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-Wl,-Bsymbolic", "-Wl,--no-undefined", "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"),
        "-lpthread"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir(os.environ.get("REFSRC", "/root/reference/src")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"])


def main(argv):
    lib = build_library(verbose="-v" in argv)
    print("built", lib)
    if "--all" in argv:
        build_oracle()
        print("built oracle")


if __name__ == "__main__":
    main(sys.argv[1:])
