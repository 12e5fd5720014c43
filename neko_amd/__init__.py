"""neko_amd: MI355X (gfx950) implementation of Neko's linear-solve hot path
(ax_helm, gs_t ADD, CG/GMRES vector kernels with glsc3 reductions, Jacobi
and hsmg preconditioners) behind a C ABI (include/neko_amd.h)."""
from ._lib import NekoAmdError, lib  # noqa: F401

__all__ = ["lib", "NekoAmdError"]
