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
