"""Build recipe (TEST INFRASTRUCTURE ONLY): compiles sid_amd/mex/sid_mex_gateway.c
once per MEX entry point against the test stand-in mex.h / mex_stub.c, linked
with sid_amd/libsidproj.so, into tests/mex_stub/_build/<name>.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
VARIANTS = {
    "forwardProjectGPU": ["-DSID_MEX_FORWARD"],
    "backwardProjectGPU_new": ["-DSID_MEX_BACKWARD"],
    "forwardProjectACC": ["-DSID_MEX_FORWARD", "-DSID_MEX_CROP"],
    "backwardProjectACC_new": ["-DSID_MEX_BACKWARD", "-DSID_MEX_CROP"],
    "sidDeconvRL": ["-DSID_MEX_ISRA"],
    "sidLSDeconv": ["-DSID_MEX_LS"],
    "generate_LFM_library_GPU": ["-DSID_MEX_LIBGPU"],
    "generate_LFM_library_CPU": ["-DSID_MEX_LIBCPU"],
    "sidPSF": ["-DSID_MEX_PSF"],
}


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    libdir = os.path.join(ROOT, "sid_amd")
    for name, defs in VARIANTS.items():
        out = os.path.join(OUT, name + ".so")
        cmd = ["gcc", "-O2", "-std=c99", "-Wall", "-Werror", "-Wno-unused-function", "-fPIC", "-shared", *defs,
               "-I", HERE, "-I", os.path.join(ROOT, "include"),
               os.path.join(ROOT, "sid_amd", "mex", "sid_mex_gateway.c"), os.path.join(HERE, "mex_stub.c"),
               "-L", libdir, "-lsidproj", "-Wl,-rpath," + libdir, "-lm", "-o", out + ".tmp"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        os.replace(out + ".tmp", out)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
