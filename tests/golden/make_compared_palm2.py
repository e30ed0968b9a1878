"""Extract the data of the reference's PALM4MSA known-answer test into
tests/golden/compared_palm2.npz.

Source: misc/config_file/mat/config_compared_palm2.mat (its `data` field, a
500 x 32 float64 matrix).  The reference's test
(misc/test/src/Python/test_FaustPy.py:1230-1263 for backend 2016, :1380-1410
for backend 2020) factorizes it with splin(500, 32, 5) and normcol(32, 32, 1.0),
200 iterations, init_lambda 1, left-to-right updates, and asserts the relative
error 0.270954109668 to 4 decimal places; the number comes from
misc/test/src/C++/test_palm4MSA.cpp.in.  Data only.  Run here, where
/root/reference exists."""
import os

import numpy as np
from scipy.io import loadmat

SRC = "/root/reference/misc/config_file/mat/config_compared_palm2.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "compared_palm2.npz")


def main():
    M = np.asarray(loadmat(SRC)["data"], dtype=np.float64)
    assert M.shape == (500, 32)
    np.savez_compressed(OUT, data=M, expected_relerr=np.float64(0.270954109668), places=np.int64(4),
                        source=np.array("misc/config_file/mat/config_compared_palm2.mat"),
                        expected_from=np.array("misc/test/src/Python/test_FaustPy.py:1263,1410"))
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()
