"""Build libmcrt.so (HIP, gfx950) in-tree, and the Fortran host shim.

Usage: python -m src3_mpi_cfitsio_amd.build  (or build() from __graft_entry__).
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libmcrt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-munsafe-fp-atomics",                      # hardware FP32/FP64 global atomics
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
# kernels_fp64 must round exactly like the reference: no FMA contraction.
# The kernel units also print the compiler's per-kernel resource usage, which
# the build checks against MIN_OCCUPANCY.
RESOURCE_REMARKS = "-Rpass-analysis=kernel-resource-usage"
UNITS = [
    ("kernels_fp64.hip", ["-ffp-contract=off", RESOURCE_REMARKS]),
    ("kernels_mixed.hip", ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt", RESOURCE_REMARKS]),
    ("mcrt_abi.hip", ["-ffp-contract=off"]),
]

# Waves per SIMD the hot kernels must keep (VGPRs <= 512 / waves).  The
# resident-lane transport kernel sizes its lanes from the occupancy it gets,
# so a few more registers in a rarely taken branch would silently halve its
# lanes (235 -> 281 VGPRs once took k_transport<float> from 2 waves to 1 and
# c1 from 1.5 to 1.3 M packets/s).
MIN_OCCUPANCY = {
    "_ZN4mcrt11k_transportIfLb0EEEvNS_11DeviceModelENS_13DeviceTalliesENS_9RunParamsE": 2,
    "_ZN4mcrt11k_transportIfLb1EEEvNS_11DeviceModelENS_13DeviceTalliesENS_9RunParamsE": 2,
    "_ZN4mcrt6k_peelIdEEvNS_11DeviceModelENS_13DeviceTalliesENS_9RunParamsE": 2,
    "_ZN4mcrt11k_peel_walkIfEEvNS_11DeviceModelENS_13DeviceTalliesENS_9RunParamsENS_8PeelRaysENS_8PeelHitsEj": 4,
}


def kernel_resources(remarks):
    """{kernel: {field: value}} from the compiler's kernel-resource-usage
    remarks ("Function Name: ...", then "VGPRs: ...", "Occupancy [waves/SIMD]:
    ..." lines)."""
    out, cur = {}, None
    for line in remarks.splitlines():
        if "remark:" not in line or "kernel-resource-usage" not in line:
            continue
        body = line.split("remark:", 1)[1].split("[-Rpass-analysis", 1)[0].strip()
        key, _, val = body.partition(":")
        key, val = key.strip(), val.strip()
        if key == "Function Name":
            cur = out.setdefault(val, {})
        elif cur is not None:
            cur[key] = val
    return out


def check_occupancy(res, unit):
    for name, need in MIN_OCCUPANCY.items():
        if unit == "kernels_mixed.hip" and ("IfE" in name or "IfL" in name) and name not in res:
            raise RuntimeError(f"{name}: no resource remark (kernel renamed? update MIN_OCCUPANCY)")
        if name in res:
            got = int(res[name].get("Occupancy [waves/SIMD]", "0"))
            if got < need:
                raise RuntimeError(f"{name}: occupancy {got} waves/SIMD with {res[name].get('VGPRs')} VGPRs "
                                   f"+ {res[name].get('AGPRs')} AGPRs; the engine is sized for {need}")


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        out = _run(cmd)
        res = kernel_resources(out)
        if res:
            check_occupancy(res, src)
            with open(os.path.join(BUILD, src.replace(".hip", ".resources.txt")), "w") as f:
                for k, v in sorted(res.items()):
                    f.write(f"{k} " + " ".join(f"{a}={b}" for a, b in v.items()) + "\n")
        elif verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """Fortran iso_c_binding module + replacement disk loop (host shim)."""
    fdir = os.path.join(PKG, "fortran")
    out = os.path.join(BUILD, "fortran")
    os.makedirs(out, exist_ok=True)
    srcs = ["mcrt_mod.f90", "disk_gpu.f90"]
    for s in srcs:
        if not os.path.exists(os.path.join(fdir, s)):
            return None
    objs = []
    for s in srcs:
        o = os.path.join(out, s.replace(".f90", ".o"))
        _run([FLANG, "-O2", "-fPIC", "-c", os.path.join(fdir, s), "-o", o], cwd=out)
        objs.append(o)
    lib = os.path.join(PKG, "libmcrt_fortran.so")
    _run([FLANG, "-shared", "-o", lib] + objs + ["-L", PKG, "-lmcrt", "-Wl,-rpath,$ORIGIN"])
    return lib


def build(verbose=False):
    lib = build_hip(verbose)
    if shutil.which(FLANG) or os.path.exists(FLANG):
        build_fortran()
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
