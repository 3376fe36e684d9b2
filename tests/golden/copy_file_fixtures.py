"""Copy the reference's own file-format test data (testsuite/testdata, used by
testsuite/file/ana.cpp and fits.cpp) into tests/golden/files: arrays whose element (j, k)
is j + k, written by the reference's ANA and FITS writers (uncompressed, and ANA's crunch
compression at slice 5 for the 40 x 50 int32 array).  Run once, from the repository
root, where the reference tree exists; the tests read only the copies."""
import os
import shutil
import sys

SRC = "/root/reference/testsuite/testdata"
NAMES = ["gradient_8u_4x5.f0", "gradient_16s_4x5_le.f0", "gradient_32s_4x5_le.f0", "gradient_32f_4x5_le.f0",
         "gradient_64f_4x5_le.f0", "gradient_32s_40x50_le.f0", "gradient_32s_40x50_le.fz",
         "gradient_8u_4x5.fits", "gradient_16s_4x5.fits", "gradient_32s_4x5.fits", "gradient_32f_4x5.fits",
         "gradient_64f_4x5.fits", "gradient_32s_40x50.fits"]

if __name__ == "__main__":
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "files")
    os.makedirs(dst, exist_ok=True)
    for n in NAMES:
        shutil.copyfile(os.path.join(SRC, n), os.path.join(dst, n))
    print(f"copied {len(NAMES)} files to {dst}", file=sys.stderr)
