"""Copies (gzip-compressed) reference output files that pin the IDL plot
file and log file formats of batsrus_amd.plotfile:

  idl_mhdhyp_1d_ref.out.gz   Param/SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out
                             (ascii, 1-D, 3 parameters)
  idl_fastwave_2d_ref.out.gz Param/SHOCKTUBE/TestOutput/fastwave_2d_ref.out
                             (ascii, 2-D, non-Cartesian: nDim = -2)
  idl_b0local_real4.out.gz   Param/CORONA/B0local.out
                             (real4 binary, 3-D, no parameters)
  log_2bodyplot.log.gz       Param/2BODYPLOT/TestOutput/REFlog_n000001.log
                             (log file, StringTime 'step time')
  PARAM.in.hyp.1d.gz         Param/SHOCKTUBE/PARAM.in.hyp.1d
                             (the parameter file of mhdhyp_1d_ref.out)

They are data files of the reference's own tests, copied unchanged.
Run where the reference tree is present:
    python tests/golden/make_idl_fixtures.py
"""
import gzip
import os
import shutil

SRC = "/root/reference/Param"
HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "idl_mhdhyp_1d_ref.out.gz": "SHOCKTUBE/TestOutput/mhdhyp_1d_ref.out",
    "idl_fastwave_2d_ref.out.gz": "SHOCKTUBE/TestOutput/fastwave_2d_ref.out",
    "idl_b0local_real4.out.gz": "CORONA/B0local.out",
    "log_2bodyplot.log.gz": "2BODYPLOT/TestOutput/REFlog_n000001.log",
    "PARAM.in.hyp.1d.gz": "SHOCKTUBE/PARAM.in.hyp.1d",
}


def main():
    for dst, src in FILES.items():
        with open(os.path.join(SRC, src), "rb") as fi, \
                gzip.GzipFile(os.path.join(HERE, dst), "wb", mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
        print("wrote", dst)


if __name__ == "__main__":
    main()
