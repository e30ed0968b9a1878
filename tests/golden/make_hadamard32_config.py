"""Extract the reference's hierarchical-factorization test configuration
misc/config_file/mat/config_HADAMARD_32.mat (used by its
MATFILE_FACT_HADAMARD test, misc/test/CMakeLists.txt:339, which expects
lambda = 5930 +- 30) into tests/golden/hadamard32_config.json.  Data only:
the constraints, iteration counts and update way, plus the expected band
from the CMake test line.  Run here, where /root/reference exists."""
import json
import os

from scipy.io import loadmat

SRC = "/root/reference/misc/config_file/mat/config_HADAMARD_32.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hadamard32_config.json")


def main():
    p = loadmat(SRC)["params"][0, 0]
    cons = p["cons"]

    def cons_row(r):
        out = []
        for c in range(cons.shape[1]):
            x = cons[r, c]
            out.append([str(x[0, 0][0]), int(x[0, 1][0, 0]), int(x[0, 2][0, 0]), int(x[0, 3][0, 0])])
        return out

    cfg = {
        "source": "misc/config_file/mat/config_HADAMARD_32.mat",
        "nfacts": int(p["nfacts"][0, 0]),
        "fact_constraints": cons_row(0),
        "res_constraints": cons_row(1),
        "niter1": int(p["niter1"][0, 0]),
        "niter2": int(p["niter2"][0, 0]),
        "update_way": int(p["update_way"][0, 0]),
        "nrow": int(p["nrow"][0, 0]),
        "ncol": int(p["ncol"][0, 0]),
        "expected_lambda": 5930.0,
        "lambda_tolerance": 30.0,
        "expected_from": "misc/test/CMakeLists.txt:339 (MATFILE_FACT_HADAMARD_{float,double})",
    }
    with open(OUT, "w") as f:
        json.dump(cfg, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
