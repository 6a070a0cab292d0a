"""MI355X-native magnification-PDF hot path of sgl_analytic.

Public API (the reference's, sgl_analytic/__init__.py): generate_pdf_lnmu,
generate_pdf; plus generate_pdf_lnmu_batch for thousands of cosmologies, and
the reference pipeline's output helpers to_image_plane and clipped_moments.
Compute runs in libsgl_amd.so (hand-written HIP for gfx950); there is no CPU
fallback.
"""
from .api import (clipped_moments, generate_pdf, generate_pdf_lnmu, generate_pdf_lnmu_batch,
                  to_image_plane)

__all__ = ["generate_pdf_lnmu", "generate_pdf", "generate_pdf_lnmu_batch", "to_image_plane",
           "clipped_moments"]
__version__ = "0.1.0"
