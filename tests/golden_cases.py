"""Loader of the committed golden fixtures (tests/golden/pade_fixtures.npz)."""
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pade_fixtures.npz")
KINDS = ["cd06", "cd10", "cf90"]
OPERS = ["d1", "d2", "filter"]


def cases():
    z = np.load(PATH, allow_pickle=False)
    keys = sorted({k.split("_")[0] for k in z.files})
    out = []
    for k in keys:
        m = z[k + "_meta"]
        out.append(dict(name=k, kind=KINDS[m[0]], which=OPERS[m[1]], n=int(m[2]), periodic=bool(m[3]),
                        bc=(int(m[4]), int(m[5])), axis=int(m[6]), dx=float(z[k + "_dx"][0]),
                        f=np.asfortranarray(z[k + "_f"]), y=z[k + "_y"]))
    return out
