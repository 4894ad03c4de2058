"""Pair histograms of xi(r_par, r_perp) from several data files (drop-in for
the reference 2pla_multiple_data.py:49-113: same flags, inputs and outputs).

    python 2pla_multiple_data.py --gpu [--precision fp64|mixed] [--counts] [--verbose]

Every ``data_dir/data*.npy`` is a separate problem, as in the reference: the
files are split over the ranks, and for each file the cut angle comes from
that file's own minimum comoving distance, its pixels are computed with
neighbours from the same file only (pairs that straddle two files are not
formed), and ``corr_dir/2d_histogram_pixel_<pixel>.npy`` is written for each
of its pixels.  A pixel present in two files is written twice; the later file
wins.  Files are taken in sorted name order (the reference's order is the
file system's listing order, which is unspecified), so reruns are identical.
"""
import argparse
import os
import sys

import numpy as np

import parameters as P
from forest_class import quasar  # noqa: F401  (needed to unpickle data*.npy)
from lya2pcf_amd import drivers


def main(argv=None):
    comm = drivers.mpi_comm()
    rank = comm.Get_rank() if comm else 0
    size = comm.Get_size() if comm else 1
    os.environ.setdefault("HIP_VISIBLE_DEVICES", str(rank % P.number_of_cuda_devices + P.cuda_device_first_number))

    ap = argparse.ArgumentParser(description="Histograms of w and w*delta, one problem per data file.")
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--cpu", action="store_true", help="reference numba CPU module (correlation_procedures_cpu)")
    g.add_argument("--gpu", action="store_true", help="MI355X HIP kernels")
    ap.add_argument("--verbose", action="store_true", help="stop after two pixels of each file")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--counts", action="store_true", help="also write per-bin pair counts")
    args = ap.parse_args(argv)

    os.makedirs(P.corr_dir, exist_ok=True)
    log = open(P.corr_dir + "thread_%d_of_%d.log" % (rank, size), "w+")
    if args.cpu:
        try:
            import correlation_procedures_cpu as correlations
        except ImportError as e:
            sys.exit("--cpu needs the reference correlation_procedures_cpu module (numba): %s" % e)
    else:
        import correlation_procedures_hip as correlations
    log.write("\nComputing 2 point correlation with \nrt_max = %s\nrp_max = %s\npixels in t = %s\npixels in p = %s"
              % (P.rtmax, P.rpmax, P.numpix_rt, P.numpix_rp))
    log.flush()
    shape_hist = (P.numpix_rp, P.numpix_rt)

    for datafile in drivers.files_of_rank(P.data_dir, rank, size):
        data = np.load(datafile, allow_pickle=True).item()
        pixels = np.array(list(data.keys()))
        print("I am node number:", rank, "I have the pixels", pixels)
        dmin, angmax = drivers.cut_angle(data, P.rtmax)
        print("Minimum comoving distance to a forest (Mpc/h):", dmin)
        print("Maximum angle between pairs of skewers that are used (rad):", angmax)
        if args.cpu:
            correlations.init(data, log, shape_hist, angmax)
        else:
            correlations.init(data, log, shape_hist, angmax, precision=args.precision, counts=args.counts)
        log.write("\nThis process computes %d pixels, which go from %s to %s"
                  % (len(pixels), pixels[0], pixels[-1]))
        for k, pixel in enumerate(pixels):
            log.write("\nComputing pixel %s, completed %d%%" % (pixel, int(k / len(pixels) * 100)))
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
