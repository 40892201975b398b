"""Known-answer vectors for the line transforms, independent of scipy and of the oracle.

Each transform is summed straight from its definition in numpy long double (80-bit x87,
~19 significant digits) for fixed, exactly representable inputs; the results are stored
as FP64 in tests/golden/transforms.npz.  Regenerate with
    python tests/golden/make_fixtures.py
Definitions (N = line length, unnormalised, as FFTW/PittPack use them):
    DFT     X_k = sum_n x_n exp(-2 pi i n k / N)
    DCT-II  X_k = 2 sum_n x_n cos(pi k (2n+1) / (2N))          (reference DCT10)
    DST-II  X_k = 2 sum_n x_n sin(pi (k+1) (2n+1) / (2N))      (reference DST10)
    DCT-IV  X_k = 2 sum_n x_n cos(pi (2k+1) (2n+1) / (4N))     (quarter-wave extension)
    DST-IV  X_k = 2 sum_n x_n sin(pi (2k+1) (2n+1) / (4N))
"""
import os

import numpy as np

LD = np.longdouble
PI = LD("3.14159265358979323846264338327950288")


def inputs(n):
    # exactly representable: small integers over a power of two
    k = np.arange(n)
    return ((k * 37 + 11) % 29 - 14).astype(np.float64) / 16.0


def dft(x):
    n = len(x)
    xs = x.astype(LD)
    j = np.arange(n, dtype=LD)
    re = np.array([np.sum(xs * np.cos(2 * PI * j * k / n)) for k in range(n)])
    im = np.array([-np.sum(xs * np.sin(2 * PI * j * k / n)) for k in range(n)])
    return re.astype(np.float64), im.astype(np.float64)


def trig(x, kind):
    n = len(x)
    xs = x.astype(LD)
    j = np.arange(n, dtype=LD)
    out = []
    for k in range(n):
        if kind == "dct2":
            out.append(2 * np.sum(xs * np.cos(PI * k * (2 * j + 1) / (2 * n))))
        elif kind == "dst2":
            out.append(2 * np.sum(xs * np.sin(PI * (k + 1) * (2 * j + 1) / (2 * n))))
        elif kind == "dct4":
            out.append(2 * np.sum(xs * np.cos(PI * (2 * k + 1) * (2 * j + 1) / (4 * n))))
        elif kind == "dst4":
            out.append(2 * np.sum(xs * np.sin(PI * (2 * k + 1) * (2 * j + 1) / (4 * n))))
    return np.array(out).astype(np.float64)


def main():
    data = {}
    for n in (8, 12, 16, 17, 24, 48, 64, 96):
        x = inputs(n)
        data[f"x_{n}"] = x
        data[f"dft_re_{n}"], data[f"dft_im_{n}"] = dft(x)
        for kind in ("dct2", "dst2", "dct4", "dst4"):
            data[f"{kind}_{n}"] = trig(x, kind)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "transforms.npz")
    np.savez(out, **data)
    print(out, len(data), "arrays")


if __name__ == "__main__":
    main()
