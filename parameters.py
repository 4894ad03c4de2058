"""Run configuration for the lya2pcf drop-in drivers (2pla.py, distortion.py).

Every name the reference drivers star-import is defined here with the
reference's meaning (reference: parameters.py:7-65), so an existing
parameters.py can be dropped in unchanged.  The MI355X path reads only the
binning, the neighbour cut and the output directory; kernel tunables (tile
size, LDS histogram layout) live in the autotuner cache, not here.
"""
import numpy as np

# Input / output directories.
data_dir = './deltas_lya2pcf/'
corr_dir = './outputs/test/'

# FITS column holding the deltas (used by the ingest only).
delta_key = "DELTA"

# Separation binning, Mpc/h.  Histograms have shape (numpix_rp, numpix_rt)
# and cover r_parallel in [0, rpmax) and r_perp in [0, rtmax).
bin_size_r = 4
rmax = 200
rpmax = rmax
rtmax = rmax
numpix_rp = rpmax // bin_size_r
numpix_rt = rtmax // bin_size_r

# Three-point binning (kept for name compatibility; not used by the 2-point path).
numpix_mu = 10
numpix_theta = 10
numpix_r = rmax // bin_size_r

# Physical constants.
lambdaa = 1215.67           # Lyman-alpha rest wavelength, Angstrom
la = np.log10(lambdaa)
c = 299792.458              # km/s
halfpi = np.pi / 2

# Forest weights: w *= ((1+z)/(1+z_ref))**(gamma/2) at ingest.
gamma = 3.8
gammaovertwo = gamma / 2
z_ref = 2.25

# HEALPix resolution of the data#.npy pixel keys.
nside = 32
# Small-angle threshold of the CPU formula (2 arcsec, with pi ~ 3.14159 as in the reference).
chiquito = 2. / 3600. * 3.14159 / 180.

# Flat LCDM used for comoving distances.
Omm = 0.3153
OmDE = 1. - Omm
d_H0 = c / 100              # Mpc/h
zmin = 0
zmax = 10
nz = 10000

# Launch shapes of the reference CUDA kernels (name compatibility only).
threads_per_block = (8, 32, 4)
threads_per_block_2 = (32, 32, 1)
max_threads = 1024

# Longest forest in the data set (the ingest rewrites this line).
max_lenght = np.int32(208)

# Neighbour cap of the reference distortion scratch (the MI355X path sizes
# its scratch from the data instead and ignores this).
number_of_neighs = 80

# Device mapping for multi-GPU runs: rank r uses device r % number_of_cuda_devices + first.
number_of_cuda_devices = 1
cuda_device_first_number = 0
