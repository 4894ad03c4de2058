"""xi procedures for MI355X, selected by ``2pla.py --gpu``.  Same module API as
the reference's correlation_procedures_pycuda (init, two_point_per_pixel)."""
from lya2pcf_amd.procedures import init, two_point_per_pixel, pair_counts, pixel_pairs  # noqa: F401
