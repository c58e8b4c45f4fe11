"""Build the in-tree libraries: libpmpic_hip.so (hipcc, gfx950) and, when amdflang is present,
the Fortran host side (component modules and the model-core stand-in driver pmpic_driver)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build_hip(jobs=5, verbose=False):
    """Compile every HIP translation unit for gfx950 and link libpmpic_hip.so (always rebuilds)."""
    csrc = os.path.join(HERE, "csrc")
    subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    out = subprocess.run(["make", "-s", f"-j{jobs}", "-C", csrc], capture_output=not verbose, text=True)
    if out.returncode != 0:
        raise RuntimeError("HIP build failed:\n" + (out.stdout or "") + (out.stderr or ""))
    return os.path.join(HERE, "libpmpic_hip.so")


def build_fortran(verbose=False):
    """Build the Fortran component shim against the C ABI (skipped when amdflang is absent)."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(os.path.join(fdir, "Makefile")) or shutil.which("amdflang") is None:
        return None
    out = subprocess.run(["make", "-s", "-B", "-C", fdir], capture_output=not verbose, text=True)
    if out.returncode != 0:
        raise RuntimeError("Fortran shim build failed:\n" + (out.stdout or "") + (out.stderr or ""))
    return fdir


def build(verbose=False):
    lib = build_hip(verbose=verbose)
    build_fortran(verbose=verbose)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
