"""DESI-layout delta files to data#.npy (drop-in for the reference
delta_reader.py: same flags and outputs).

    python delta_reader.py --delta-dir DIR [--data-dir DIR] [--split-number N]

Reads every ``DIR/*.fits.gz`` (LAMBDA, DELTA, WEIGHT images and a METADATA
table, see lya2pcf_amd.ingest), builds one ``forest_class.quasar`` per forest,
groups them by HEALPix pixel, attaches each forest's neighbour names and
pixels under its own cut angle, and writes ``data1.npy .. dataN.npy`` (the
sorted pixels split in N parts), plus the text files ``sizes`` and
``neighbors`` in the working directory.  Files are read in sorted name order
(the reference maps a process pool over the file system's listing order).
The reference also rewrites ``max_lenght`` in parameters.py for its padded
GPU layout; the HIP kernels use ragged arrays, so the value is printed and
parameters.py is left alone.
"""
import argparse
import glob
import os
import warnings
from multiprocessing import Pool

import numpy as np

import parameters as P
from forest_class import quasar
from lya2pcf_amd import ingest


def _read(path):
    print("Extracting from file ", path)
    return ingest.records_desi(path, ingest.Settings.from_parameters(P))


def main(argv=None):
    ap = argparse.ArgumentParser(formatter_class=argparse.ArgumentDefaultsHelpFormatter,
                                 description="Takes delta files by picca and stores data in data.npy.")
    ap.add_argument("--delta-dir", type=str, required=True, help="Path to the delta files.")
    ap.add_argument("--data-dir", type=str, default=P.data_dir, help="Directory where the data will be stored.")
    ap.add_argument("--split-number", type=int, default=1, help="Number of files to split the data.")
    args = ap.parse_args(argv)

    if os.path.exists(args.data_dir):
        warnings.warn("The output delta directory already exists. This procedure might mix deltas from a "
                      "different run.")
    os.makedirs(args.data_dir, exist_ok=True)
    files = sorted(glob.glob(args.delta_dir + "/*.fits.gz"))
    if not files:
        print("No delta files in directory " + args.delta_dir)
    if len(files) > 1:
        with Pool(min(len(files), os.cpu_count() or 1, 16)) as pool:
            per_file = pool.map(_read, files)
    else:
        per_file = [_read(f) for f in files]
    dist = ingest.Settings.from_parameters(P).distance()
    data, sizes = {}, []
    for recs in per_file:
        for q in ingest.quasars(recs, quasar, dist):
            data.setdefault(q.pix, []).append(q)
            sizes.append(len(q.dc))
    del per_file
    if not data:
        raise NameError("Empty data. Did you write correctly the input directory?")
    min_distance = min(q.dc[0] for pix in data for q in data[pix])
    max_lenght = max(sizes)
    print("Minimum comoving distance to a forest (Mpc/h):", min_distance)
    print("Maximum angle between pairs of skewers that are used (rad):", 2 * np.arcsin(0.5 * P.rtmax / min_distance))

    neighbors = ingest.neighbour_names(data, P.rtmax)
    np.savetxt("sizes", sizes)
    np.savetxt("neighbors", neighbors)

    pixels = sorted(data.keys())
    for i, subset in enumerate(np.array_split(pixels, args.split_number), start=1):
        np.save(args.data_dir + "/data" + str(i), {int(x): data[x] for x in subset})
    print("The largest forest has ", max_lenght, " data points.")
    print("The number of forests is:", len(sizes))


if __name__ == "__main__":
    main()
