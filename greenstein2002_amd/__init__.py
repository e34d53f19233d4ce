"""MI355X-native release-unit Monte Carlo of the Greenstein-Winslow 2002
canine ventricular myocyte model.

The FRU hot path (simfru_local / fru_rates_local / fcn_fru / MT19937 and the
flux and average reductions) runs as hand-written HIP kernels for gfx950 in
``libgrw_fru.so`` behind the C ABI of ``include/grw_fru.h``; ``FruCell`` mirrors
the reference's master-side interface over it.
"""
from ._lib import GrwError
from .fru import FruCell

__all__ = ["FruCell", "GrwError"]


def __getattr__(name):  # the whole-cell and batch layers load on first use
    if name in ("Cell", "Protocol"):
        from . import cell

        return getattr(cell, name)
    if name in ("Batch", "CellSpec"):
        from . import batch

        return getattr(batch, name)
    if name == "run_caf":
        from . import output

        return output.run_caf
    raise AttributeError(name)
