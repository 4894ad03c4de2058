"""Helpers shared by the drop-in driver scripts (2pla*.py, distortion*.py)."""
import glob

import numpy as np


def mpi_comm():
    """MPI.COMM_WORLD when mpi4py is importable, else None (one rank)."""
    try:
        from mpi4py import MPI
        return MPI.COMM_WORLD
    except ImportError:
        return None


def data_files(data_dir):
    """The ``data*.npy`` files of a directory (reference 2pla_multiple_data.py:50),
    in sorted name order."""
    files = sorted(glob.glob(data_dir + "/data*.npy"))
    if not files:
        raise FileNotFoundError("no data*.npy files in " + data_dir)
    return files


def files_of_rank(data_dir, rank, size):
    """This rank's share of the data files (np.array_split, as the reference)."""
    return [str(f) for f in np.array_split(np.array(data_files(data_dir)), size)[rank]]


def cut_angle(data, rtmax):
    """(dmin, angmax = 2 asin(rtmax / (2 dmin))) over the forests of one data dict
    (reference 2pla.py:64-73)."""
    dmin = [f.dc[0] for pix in data for f in data[pix]]
    if not dmin:
        raise NameError("Empty data. Did you write correctly the input directory?")
    dmin = min(dmin)
    return dmin, 2 * np.arcsin(0.5 * rtmax / dmin)
