"""Build recipe for the in-tree native code (run on any host; hipcc
cross-compiles gfx950 without a GPU).

    libpsivox.so   HIP kernels + C ABI (include/psivox.h), hipcc --offload-arch=gfx950
    PSI<ext>.so    CPython module "PSI" (C), linked to libpsivox.so with an
                   $ORIGIN rpath so the pair travels together.
    atomic_rate    measurement program: FP64 atomic throughput of the
                   deposit pattern (csrc/atomic_rate.hip)

python -m pypsi_amd.build   builds both.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBPSIVOX = os.path.join(HERE, "libpsivox.so")
EXT = os.path.join(HERE, "PSI" + sysconfig.get_config_var("EXT_SUFFIX"))
ATOMIC_RATE = os.path.join(HERE, "atomic_rate")
ARCH = os.environ.get("PSIVOX_ARCH", "gfx950")

HIP_SOURCES = ["psivox.hip"]
HIP_HEADERS = ["psivox_geom.h", "psivox_plain.h", "psivox_refine.h", "psivox_fast.h", "psivox_adjoint.h"]

HIPCC_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 VALU forms (v_pk_*_f32) stay off: some of their forms have
    # given wrong results on these GPUs without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build_hip(force: bool = True) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    cmd = ["hipcc"] + HIPCC_FLAGS + ["-I" + os.path.join(ROOT, "include"), "-o", LIBPSIVOX] + srcs
    _run(cmd)
    return LIBPSIVOX


def build_module() -> str:
    import numpy as np
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-Wall", "-Wno-missing-field-initializers",
           "-I" + py_inc, "-I" + np.get_include(), "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "psimodule.c"), "-L" + HERE, "-lpsivox", "-Wl,-rpath,$ORIGIN", "-o", EXT]
    _run(cmd)
    return EXT


def build_probe() -> str:
    flags = [f for f in HIPCC_FLAGS if f not in ("-fPIC", "-shared")]
    _run(["hipcc"] + flags + ["-o", ATOMIC_RATE, os.path.join(CSRC, "atomic_rate.hip")])
    return ATOMIC_RATE


def build_all() -> None:
    build_hip()
    build_module()
    build_probe()


if __name__ == "__main__":
    build_all()
    print("built", LIBPSIVOX, EXT)
