"""Write the small fixtures of tests/golden from the reference's data files (data only; no reference
code is imported or run):

  planck18_pairs.npz   (z, D_L) of the simulated segments of the reference, computed there with
                       astropy Planck18 (data/segments/two_parameters.npz, family.npz members
                       true_z, true_luminosity_distance_mpc)
  segments_2p.npz      rows 0-1 of data/segments/two_parameters.npz in the segment file format
  s_grid.npz           the 53 store nodes (data/tables/s_grid.npz)

    python tests/golden/make_fixtures.py /root/reference
"""
import os
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
out = os.path.dirname(os.path.abspath(__file__))
z, dl = [], []
for name in ("two_parameters", "family"):
    with np.load(os.path.join(ref, "data", "segments", name + ".npz")) as f:
        z.append(f["true_z"])
        dl.append(f["true_luminosity_distance_mpc"])
np.savez(os.path.join(out, "planck18_pairs.npz"), z=np.concatenate(z), dl=np.concatenate(dl))
with np.load(os.path.join(ref, "data", "segments", "two_parameters.npz")) as f:
    np.savez(os.path.join(out, "segments_2p.npz"), d_H=f["d_H"][:2], d_L=f["d_L"][:2], f=f["f"], nu=f["nu"],
             sig2=f["sig2"], T=f["T"], fmin=f["fmin"], fmax=f["fmax"])
with np.load(os.path.join(ref, "data", "tables", "s_grid.npz")) as f:
    np.savez(os.path.join(out, "s_grid.npz"), s_nodes=f["s_nodes"])
print("fixtures written to", out)
