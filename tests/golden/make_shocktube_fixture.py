"""Writes tests/golden/mhdhyp_1d_ref.npz from the reference's committed
regression output Param/SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out (data only:
the numeric columns of the IDL ascii plot file).

That run (Param/SHOCKTUBE/PARAM.in.hyp.1d, Makefile.test:303-323) is a 1-D
Brio-Wu shock tube with gamma=2 on 1024 cells (16 blocks of 64) over
[-128, 128], first-order Rusanov, one stage, fixed dt = 0.05, float outer
boundaries, 512 steps to t = 25.6; the reference checks it with
DiffNum.pl -r=1e-5 -a=1e-11.

Run where the reference tree is present (it reads the reference's output):
    python tests/golden/make_shocktube_fixture.py
"""
import os
import sys

import numpy as np

SRC = "/root/reference/Param/SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "mhdhyp_1d_ref.npz")


def main():
    with open(SRC) as f:
        lines = f.read().splitlines()
    head = lines[1].split()
    nstep, time = int(head[0]), float(head[1])
    ncell = int(lines[2].split()[0])
    params = [float(v) for v in lines[3].split()]
    names = lines[4].split()
    data = np.array([[float(v) for v in ln.split()] for ln in lines[5:5 + ncell]])
    cols = names[:data.shape[1]]
    np.savez_compressed(DST, nstep=nstep, time=time, gamma=params[0],
                        names=np.array(cols), data=data)
    print("wrote", DST, data.shape, cols)


if __name__ == "__main__":
    sys.exit(main())
