"""Distortion matrix of the projected xi (drop-in for the reference
distortion.py: same flags, input and output).

    python distortion.py [--excluded 0.95] [--verbose]

Reads ``data_dir + 'data1.npy'``, keeps a fraction 1 - excluded of each
forest's neighbours (reference subsampling, random.seed(1)), and writes
``corr_dir + 'distortion.npy'`` = D / weight[:, None] as float64
(np_rp*np_rt, np_rp*np_rt) (reference distortion.py:92-125).  Sums are
accumulated in FP64 on the device across all pixels of the rank.
"""
import argparse
import os

import numpy as np

import parameters as P
from forest_class import quasar  # noqa: F401


def main(argv=None):
    try:
        from mpi4py import MPI
        comm = MPI.COMM_WORLD
    except ImportError:
        comm = None
    rank = comm.Get_rank() if comm else 0
    size = comm.Get_size() if comm else 1
    os.environ.setdefault("HIP_VISIBLE_DEVICES", str(rank % P.number_of_cuda_devices + P.cuda_device_first_number))
    import distortion_procedures_hip as distortion

    ap = argparse.ArgumentParser(description="Distortion matrix of the two-point correlation function.")
    ap.add_argument("--excluded", default=0.95, help="fraction of forest pairs excluded")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args(argv)

    os.makedirs(P.corr_dir, exist_ok=True)
    log = open(P.corr_dir + "thread_%d_of_%d_distortion.log" % (rank, size), "w+")
    data = np.load(P.data_dir + "data1.npy", allow_pickle=True).item()
    shape_hist = (P.numpix_rp, P.numpix_rt)
    pixels = np.array(list(data.keys()))
    mine = np.array_split(pixels, size)[rank]
    dmin = [f.dc[0] for pix in data for f in data[pix]]
    if not dmin:
        raise NameError("Empty data. Did you write correctly the input directory?")
    angmax = 2 * np.arcsin(0.5 * P.rtmax / min(dmin))
    distortion.init(data, log, shape_hist, angmax, float(args.excluded))

    # device-side accumulation over the rank's pixels (same subsample sequence as
    # calling distortion_per_pixel pixel by pixel)
    D, wA = distortion.distortion_of_pixels(mine[:2] if args.verbose else mine, log)
    if comm is not None and size > 1:
        D = comm.reduce(D)
        wA = comm.reduce(wA)
    if rank == 0:
        with np.errstate(divide="ignore", invalid="ignore"):
            np.save(P.corr_dir + "distortion", D / wA[:, None])
    print("Finished distortion computation.")


if __name__ == "__main__":
    main()
