"""Distortion matrix from several data files (drop-in for the reference
distortion_multiple_data.py:49-136: same flags, inputs and output).

    python distortion_multiple_data.py [--excluded 0.95] [--verbose]

As in the reference, the ``data_dir/data*.npy`` files are split over the
ranks and every file is a separate problem: its own cut angle, neighbours
from the same file only, and every pixel of the file computed by the rank
that holds the file.  The sums of D and of the bin weights run over all the
rank's files (FP64), are summed over ranks, and rank 0 writes
``corr_dir/distortion.npy`` = D / weight[:, None].  The neighbour subsample
draws from one random stream seeded once with 1 (as the reference module
does at import), continuing from file to file in sorted name order.
"""
import argparse
import os

import numpy as np

import parameters as P
from forest_class import quasar  # noqa: F401
from lya2pcf_amd import drivers


def main(argv=None):
    comm = drivers.mpi_comm()
    rank = comm.Get_rank() if comm else 0
    size = comm.Get_size() if comm else 1
    os.environ.setdefault("HIP_VISIBLE_DEVICES", str(rank % P.number_of_cuda_devices + P.cuda_device_first_number))
    import distortion_procedures_hip as distortion

    ap = argparse.ArgumentParser(description="Distortion matrix, one problem per data file.")
    ap.add_argument("--excluded", default=0.95, help="fraction of forest pairs excluded")
    ap.add_argument("--verbose", action="store_true", help="stop after two pixels of each file")
    args = ap.parse_args(argv)

    os.makedirs(P.corr_dir, exist_ok=True)
    log = open(P.corr_dir + "thread_%d_of_%d_distortion.log" % (rank, size), "w+")
    shape_hist = (P.numpix_rp, P.numpix_rt)
    nb = int(np.prod(shape_hist))
    D = np.zeros((nb, nb))
    wA = np.zeros(nb)
    for datafile in drivers.files_of_rank(P.data_dir, rank, size):
        print("Loading extracted file.")
        data = np.load(datafile, allow_pickle=True).item()
        pixels = np.array(list(data.keys()))
        dmin, angmax = drivers.cut_angle(data, P.rtmax)
        print("Minimum comoving distance to a forest (Mpc/h):", dmin)
        print("Maximum angle between pairs of skewers that are used (rad):", angmax)
        distortion.init(data, log, shape_hist, angmax, float(args.excluded))
        log.write("\nNumber of non-empty healpix pixels: %d" % len(pixels))
        stop = 2 if args.verbose else None
        d, w = distortion.distortion_of_pixels(pixels[:stop], log)
        D += d
        wA += w
        print("Finished data file " + datafile)
    if comm is not None and size > 1:
        D = comm.reduce(D)
        wA = comm.reduce(wA)
    if rank == 0:
        with np.errstate(divide="ignore", invalid="ignore"):
            np.save(P.corr_dir + "distortion", D / wA[:, None])
    print("Finished distortion computation.")


if __name__ == "__main__":
    main()
