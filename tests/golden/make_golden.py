"""Writes tests/golden/fft_golden.npz: inputs and numpy/pocketfft float64 outputs.

Run once in the development container (numpy only; no reference code is imported):
    python tests/golden/make_golden.py
Each case stores <name>_in, <name>_out and <name>_meta = [rank, n0, n1, n2, batch, sign, norm].
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NORMS = ["none", "forward", "backward", "orthogonal"]


def case(shape, batch, sign, norm, seed):
    g = np.random.default_rng(seed)
    n = int(np.prod(shape)) * batch
    x = (g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)).reshape((batch,) + tuple(shape))
    axes = tuple(range(1, len(shape) + 1))
    y = np.fft.fftn(x, axes=axes) if sign == -1 else np.fft.ifftn(x, axes=axes) * np.prod(shape)
    t = np.prod(shape)
    s = {"none": 1.0, "forward": 1 / t if sign == -1 else 1.0,
         "backward": 1 / t if sign == 1 else 1.0, "orthogonal": 1 / np.sqrt(t)}[norm]
    meta = np.array([len(shape)] + list(shape) + [1] * (3 - len(shape)) + [batch, sign, NORMS.index(norm)],
                    dtype=np.int64)
    return x.reshape(-1), (y * s).reshape(-1), meta


def main():
    cases = {
        "c1_pow2_4096": ((4096,), 2, -1, "none"),
        "c2_mixed_2520": ((2520,), 2, 1, "orthogonal"),
        "c2_pow3_6561": ((6561,), 1, -1, "orthogonal"),
        "c2_prime_10007": ((10007,), 1, -1, "none"),
        "prime_257": ((257,), 3, 1, "backward"),
        "smooth_240": ((240,), 4, -1, "forward"),
        "md_3d_16_12_10": ((16, 12, 10), 2, -1, "forward"),
        "md_2d_64_48": ((64, 48), 1, 1, "none"),
        "md_2d_11_3": ((11, 3), 1, -1, "none"),
        "tiny_8": ((8,), 2, -1, "orthogonal"),
    }
    out = {}
    for i, (name, (shape, batch, sign, norm)) in enumerate(sorted(cases.items())):
        x, y, meta = case(shape, batch, sign, norm, 20260100 + i)
        out[name + "_in"] = x
        out[name + "_out"] = y
        out[name + "_meta"] = meta
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
