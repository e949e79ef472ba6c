"""
coulomb_kmc_amd: MI355X-native (gfx950) FMM-KMC propose/accept, a drop-in for the
propose/accept path of ppmd/coulomb_kmc.
"""

from .common import BCType, Domain, cell_offsets, spherical
from .kmc_fmm import KMCFMM
from .kmc_inject_extract import DiscoverInjectExtract, InjectorExtractor

__all__ = ["KMCFMM", "BCType", "Domain", "cell_offsets", "spherical", "DiscoverInjectExtract", "InjectorExtractor"]
