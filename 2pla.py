"""Pair histograms of xi(r_par, r_perp) per HEALPix pixel (drop-in for the
reference 2pla.py: same flags, inputs and output files).

    python 2pla.py --gpu [--precision fp64|mixed] [--counts] [--verbose]
    python 2pla.py --cpu          (the reference numba module, if installed)

Reads ``data_dir + 'data1.npy'`` and writes
``corr_dir + '2d_histogram_pixel_<pixel>.npy'`` = float64 (2, np_rp, np_rt) with
(sum w1 w2, sum dw1 dw2) for the pairs whose forest 1 lies in <pixel>
(reference 2pla.py:52-140).  --counts also writes
``2d_paircounts_pixel_<pixel>.npy`` (a name post-processing does not glob).
Multi-rank runs use mpi4py when it is importable, one GPU per rank.
"""
import argparse
import os
import sys

import numpy as np

from parameters import *  # noqa: F401,F403
import parameters as P
from forest_class import quasar  # noqa: F401  (needed to unpickle data1.npy)


def _comm():
    try:
        from mpi4py import MPI
        return MPI.COMM_WORLD
    except ImportError:
        return None


def main(argv=None):
    comm = _comm()
    rank = comm.Get_rank() if comm else 0
    size = comm.Get_size() if comm else 1
    dev = rank % P.number_of_cuda_devices + P.cuda_device_first_number
    os.environ.setdefault("HIP_VISIBLE_DEVICES", str(dev))

    ap = argparse.ArgumentParser(description="Histograms of w and w*delta for the two-point correlation.")
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--cpu", action="store_true", help="reference numba CPU module (correlation_procedures_cpu)")
    g.add_argument("--gpu", action="store_true", help="MI355X HIP kernels")
    ap.add_argument("--verbose", action="store_true", help="stop after two pixels")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--counts", action="store_true", help="also write per-bin pair counts")
    args = ap.parse_args(argv)

    os.makedirs(P.corr_dir, exist_ok=True)
    log = open(P.corr_dir + "thread_%d_of_%d.log" % (rank, size), "w+")
    print("Loading extracted file.")
    data = np.load(P.data_dir + "data1.npy", allow_pickle=True).item()
    dmin = [f.dc[0] for pix in data for f in data[pix]]
    if not dmin:
        raise NameError("Empty data. Did you write correctly the input directory?")
    angmax = 2 * np.arcsin(0.5 * P.rtmax / min(dmin))
    shape_hist = (P.numpix_rp, P.numpix_rt)
    print("Minimum comoving distance to a forest (Mpc/h):", min(dmin))
    print("Maximum angle between pairs of skewers that are used (rad):", angmax)
    pixels = np.array(list(data.keys()))
    mine = np.array_split(pixels, size)[rank]
    log.write("\nNumber of non-empty healpix pixels: %d" % len(pixels))

    if args.cpu:
        try:
            import correlation_procedures_cpu as correlations
        except ImportError as e:
            sys.exit("--cpu needs the reference correlation_procedures_cpu module (numba): %s" % e)
        correlations.init(data, log, shape_hist, angmax)
    else:
        import correlation_procedures_hip as correlations
        correlations.init(data, log, shape_hist, angmax, compute_pixels=list(mine), precision=args.precision,
                          counts=args.counts)

    for k, pixel in enumerate(mine):
        log.write("\nComputing pixel %s, completed %d%%" % (pixel, int(k / max(1, len(mine)) * 100)))
        log.flush()
        histo = correlations.two_point_per_pixel(pixel)
        np.save(P.corr_dir + "2d_histogram_pixel_" + str(pixel), histo)
        if args.counts and args.gpu:
            np.save(P.corr_dir + "2d_paircounts_pixel_" + str(pixel), correlations.pair_counts(pixel))
        if args.verbose and k > 0:
            print("Exiting early due to --verbose option.")
            break
    print("Finished correlation computation.")


if __name__ == "__main__":
    main()
