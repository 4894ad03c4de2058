"""Distortion procedures for MI355X, used by ``distortion.py``.  Same module API
as the reference's distortion_procedures_pycuda (init, distortion_per_pixel)."""
from lya2pcf_amd.distortion import init, distortion_per_pixel  # noqa: F401
