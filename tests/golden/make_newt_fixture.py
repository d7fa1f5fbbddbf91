"""Write tests/golden/newt_poiseuille.npz from the reference-held output files.

The reference ships a saved state of its Newtonian channel example
(higflow/example2d_Newt/output/newt.save._dp{p,u0,u1}, step 1000, t = 1.0;
written by pdomain.c:1328-1385 as DIM+1 float64 records (x, y, value) in
global-id order).  This script decodes those data files with ``struct`` and
stores the records as a compressed fixture.  It runs only where the reference
tree is mounted; the committed .npz is what the tests read.
"""
import os
import struct
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "higflow", "example2d_Newt", "output")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "newt_poiseuille.npz")


def records(name):
    raw = open(os.path.join(SRC, name), "rb").read()
    vals = np.array(struct.unpack("<%dd" % (len(raw) // 8), raw), dtype=np.float64)
    return vals.reshape(-1, 3)


def main():
    data = {k: records("newt.save._dp" + k) for k in ("p", "u0", "u1")}
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, {k: v.shape for k, v in data.items()})


if __name__ == "__main__":
    main()
