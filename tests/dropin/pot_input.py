"""Writes an input file for tests/dropin/pot_main.cpp: n masses and m points
(x, y, z ~ N(0.5, 0.1), masses ~ U(0,1)/n, seed 7) on a g^3 grid with near
range K.  Used for the potential-flow path's kernel-trace profile.

usage: python3 pot_input.py <out.bin> [n] [g] [K]
"""
import sys

import numpy as np


def main():
    path = sys.argv[1]
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 1 << 20
    g = int(sys.argv[3]) if len(sys.argv) > 3 else 128
    K = int(sys.argv[4]) if len(sys.argv) > 4 else 2
    m = n
    rng = np.random.default_rng(7)
    mp = np.zeros((n, 4), np.float32)
    mp[:, :3] = (0.5 + 0.1 * rng.standard_normal((n, 3))).astype(np.float32)
    mass = rng.random(n) / n
    ev = np.zeros((m, 4), np.float32)
    ev[:, :3] = (0.5 + 0.1 * rng.standard_normal((m, 3))).astype(np.float32)
    with open(path, "wb") as f:
        f.write(np.array([n, m, g], np.uint32).tobytes())
        f.write(np.array([K], np.int32).tobytes())
        for a in (mp, mass, ev):
            f.write(np.ascontiguousarray(a).tobytes())
    print(f"wrote {path}: n = m = {n}, g = {g}, K = {K}")


if __name__ == "__main__":
    main()
