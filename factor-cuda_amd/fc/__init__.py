"""fc: cross-sectional factor analysis on MI355X (gfx950).

Drop-in for the reference package's public API (fc/__init__.py:5-29):
cross_sectional_rank, factor_plane, parameter_scan, rolling_ic,
factor_corr, stock_corr. GPU work runs in libfc_hip.so through the
factor_cuda_pybind / factor_corr_pybind modules built into
factor-cuda_amd/build; the bindings load lazily, so importing fc needs no GPU.
"""
from .cross_sectional_rank import cross_sectional_rank, factor_plane
from .correlation import factor_corr, stock_corr, set_corr_precision, get_corr_precision
from .rolling_ic import rolling_ic
from .parameter_scan import parameter_scan
from ._workspace import _clear_all as _clear_workspaces


def clear_workspaces():
    """Release every cached device workspace held by the adapter (idempotent).
    The next call re-creates them; never required for correctness."""
    _clear_workspaces()


__version__ = "1.1.1"
# set_corr_precision / clear_workspaces are module attributes but not part of
# the locked public surface (__all__ mirrors the reference's).
__all__ = ["cross_sectional_rank", "factor_plane", "factor_corr", "stock_corr",
           "rolling_ic", "parameter_scan", "__version__"]
