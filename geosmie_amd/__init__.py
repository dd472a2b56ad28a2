"""geosmie_amd: GEOSmie's Mie / LUT hot path on AMD MI355X (gfx950).

Drop-in Python surface of the reference's operator API:
  pymiecoated   Mie, MieCoeffs, mie_coeffs, mie_props, mie_S12, MultipleMie
  runoptics     rawMie, integratePSD, calculateScatVals, normalize_pmatrix
  gsf           spher_expan (generalised-spherical-function expansion), hbleg
backed by hand-written HIP kernels in libgeosmie_hip.so (C ABI in
include/geosmie.h).  There is no CPU compute path.
"""
from .mie import Mie, MieCoeffs, MieScatterProps, mie_coeffs, mie_props, mie_S12, mie_S12_pt

__all__ = ["Mie", "MieCoeffs", "MieScatterProps", "mie_coeffs", "mie_props", "mie_S12",
           "mie_S12_pt"]
