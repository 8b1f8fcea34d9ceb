"""MI355X-native GrB_mxm (dot3 + saxpy3) behind the GraphBLAS C API.

The product is the C shared library ``lib/libgraphblas_mi355x.so`` built from
``host/`` (C, the GrB_* API) and ``csrc/`` (HIP kernels for gfx950 and the
gm_* C-ABI).  ``capi`` binds it with ctypes; ``workloads`` generates the
synthetic inputs.
"""
import os

__all__ = ["capi", "workloads", "build"]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> None:
    """Compile every source of the package (always rebuilds)."""
    import subprocess
    cmd = ["make", "-B", "-j8", "-C", PKG_DIR]
    out = subprocess.run(cmd, capture_output=not verbose, text=True)
    if out.returncode != 0:
        raise RuntimeError("build failed:\n" + (out.stdout or "") + (out.stderr or ""))
