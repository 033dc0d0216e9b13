"""MI355X-native distance-driven CT projectors (deeprecon drop-in).

Layers, top down:
  pridediff_amd.deeprecon.torch.projector  the reference's Python API
      (projector2d / projector3d / projectorsp, autograd Functions, filters);
  pridediff_amd.radontorch                 the torch-op boundary (37 ops,
      same names and signatures as the reference's radontorch module);
  libdd.so                                 HIP kernels for gfx950 behind a
      C ABI (include/dd.h, include/dd3d.h).
  pridediff_amd.batched                    one-launch batched FP/adjoint for
      many instances (uniform or uneven), graph-captured steps;
  pridediff_amd.autotune                   tile / kernel-variant / grid tuner;
  pridediff_amd.fbp                        FBP as one GEMM + weighted
      backprojection, PrideDiff's data-consistency step;
  pridediff_amd.views                      view-resolved backprojection.

``install_dropin()`` registers the modules under the reference's import names
(``radontorch``, ``deeprecon``) so unmodified callers pick them up.
"""
import importlib
import sys

__all__ = ["install_dropin", "load_radontorch"]


def load_radontorch():
    """Imports the compiled op module; raises if it has not been built."""
    import torch  # noqa: F401  (libdd must share torch's HIP runtime)
    try:
        return importlib.import_module("pridediff_amd.radontorch")
    except ImportError as e:  # pragma: no cover - exercised only when unbuilt
        raise ImportError(
            "pridediff_amd.radontorch is not built; run `python -m pridediff_amd.build`"
        ) from e


def install_dropin():
    """Makes ``import radontorch`` and ``import deeprecon.torch.projector``
    resolve to this package (the reference's module names)."""
    sys.modules["radontorch"] = load_radontorch()
    dr = importlib.import_module("pridediff_amd.deeprecon")
    sys.modules["deeprecon"] = dr
    for sub in ("torch", "torch.projector", "torch.projector.projector",
                "torch.projector.function", "torch.projector.filter"):
        sys.modules[f"deeprecon.{sub}"] = importlib.import_module(f"pridediff_amd.deeprecon.{sub}")
