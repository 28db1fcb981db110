"""Write fftw_dct_kat.npz: FFTW-generated known-answer vectors of REDFT10
(DCT-II) and REDFT01 (DCT-III), unnormalised, input linspace(0, N-1, N),
copied from the reference data file that ships with scipy
(scipy/fftpack/tests/fftw_double_ref.npz)."""
import os

import numpy as np
import scipy

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(os.path.dirname(scipy.__file__), "fftpack", "tests", "fftw_double_ref.npz")
    d = np.load(src)
    out = {}
    for key in d.files:
        parts = key.split("_")
        if len(parts) == 3 and parts[0] == "dct" and parts[1] in ("2", "3") and int(parts[2]) <= 512:
            out[key] = d[key]
    np.savez_compressed(os.path.join(HERE, "fftw_dct_kat.npz"), **out)


if __name__ == "__main__":
    main()
