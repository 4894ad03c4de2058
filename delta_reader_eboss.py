"""eBOSS-layout delta files to data1.npy (drop-in for the reference
delta_reader_eboss.py: same flags and output).

    python delta_reader_eboss.py --delta-dir DIR [--data-dir DIR]

Reads every ``DIR/*.fits.gz`` (one table HDU per forest: LOGLAM, DELTA,
WEIGHT; FIBERID, PLATE, THING_ID, RA, DEC in the header, see
lya2pcf_amd.ingest), builds one ``forest_class.quasar`` per forest and writes
``data1.npy`` = {pixel: [quasar]} with pixels in order of first appearance.
Files are read in sorted name order (the reference maps a process pool over
the file system's listing order).
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
    return ingest.records_eboss(path, ingest.Settings.from_parameters(P))


def main(argv=None):
    ap = argparse.ArgumentParser(formatter_class=argparse.ArgumentDefaultsHelpFormatter,
                                 description="Takes delta files by picca and stores data in data.npy.")
    ap.add_argument("--delta-dir", type=str, required=True, help="Path to the delta files.")
    ap.add_argument("--data-dir", type=str, default=P.data_dir, help="Directory where the data will be stored.")
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
    data = {}
    for recs in per_file:
        for q in ingest.quasars(recs, quasar, dist):
            data.setdefault(q.pix, []).append(q)
    np.save(args.data_dir + "/data1", data)


if __name__ == "__main__":
    main()
