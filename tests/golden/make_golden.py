"""Writes the committed golden fixtures of tests/golden/.

expm_su3.npz: su(3) matrices Z (traceless anti-Hermitian, norms from 1e-3 to
~1.3), coefficients c and scipy.linalg.expm(c*Z) -- the independent pin of
Luescher's Cayley-Hamilton exponential (reference src/su3/utils.cpp:145-193).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
from scipy.linalg import expm

HERE = os.path.dirname(os.path.abspath(__file__))


def random_su3_algebra(rng, scale):
    A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    Z = 0.5 * (A - A.conj().T)
    Z -= np.trace(Z) / 3 * np.eye(3)
    return scale * Z / np.linalg.norm(Z)


def main():
    rng = np.random.default_rng(20260101)
    scales = np.concatenate([np.geomspace(1e-3, 1.3, 20), rng.uniform(0.01, 0.6, 12)])
    Z = np.stack([random_su3_algebra(rng, s) for s in scales])
    coeff = rng.choice([1.0, 0.25 * 0.02, 0.02, 0.5, -0.7], size=len(Z))
    out = np.stack([expm(c * z) for c, z in zip(coeff, Z)])
    np.savez(os.path.join(HERE, "expm_su3.npz"), Z=Z, coeff=coeff, expm=out)
    print("wrote", len(Z), "expm vectors")


if __name__ == "__main__":
    main()
