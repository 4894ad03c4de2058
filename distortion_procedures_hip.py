"""Distortion procedures for MI355X, used by ``distortion.py``.  Same module API
as the reference's distortion_procedures_pycuda (init, distortion_per_pixel), plus
distortion_of_pixels (the sum over pixels, accumulated on the device)."""
from lya2pcf_amd.distortion import init, distortion_per_pixel, distortion_of_pixels  # noqa: F401
