"""Writes tests/golden/xi_tiny.npz: a 24-forest synthetic set (seed 99), its
pair list and the oracle's GPU-form histograms.  The reference ships no data
or golden vectors; this fixture freezes the generator + oracle outputs so a
change in either is caught."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from lya2pcf_amd import forests, synth  # noqa: E402


def main():
    s = synth.generate(synth.scaled(synth.CONFIGS["patch2k"], 24, seed=99, length_range=(20, 40)))
    fs = forests.ForestSet.from_survey(s)
    am = forests.angmax_from(fs, 200.0)
    off, idx = oracle.pairs(fs, am)
    c, w, d = oracle.xi_gpu_form(fs, off, idx)
    np.savez_compressed(os.path.join(HERE, "xi_tiny.npz"), pix_off=fs.pix_off, slots=np.array(fs.slots),
                        ux=fs.ux, uy=fs.uy, uz=fs.uz, ra=fs.ra, dec=fs.dec, name=fs.name, slot=fs.slot,
                        omega=fs.omega, odl2=fs.odl2, dc=fs.dc, we=fs.we, dw=fs.dw, dl=fs.dl, angmax=am,
                        nb_off=off, nb_idx=idx, cnt=c, w=w, dw_sum=d)
    print("pairs", len(idx), "binned pixel pairs", int(c.sum()))


if __name__ == "__main__":
    main()
