"""Writes tests/golden/pca_fixtures.npz: scikit-learn PCA applied exactly as
SVD_MFS.py:38-110 applies it (rank 2, channel-major flattening, V2 = V0 x V1,
B = transform + mean.V), on seeded synthetic uvw.  Needs scikit-learn
(1.7.2 here); run from the repository root.  The fixtures pin the FP64
projection of the oracle and of the GPU path without importing the
reference script.
"""
import os

import numpy as np
from sklearn.decomposition import PCA

C = 299792458.0


def case(rng, nrows, nchan, tilt):
    # A planar array tilted out of the uv plane, plus a little scatter.
    xy = rng.normal(size=(nrows, 2)) * rng.uniform(500, 5000, size=(1, 2))
    w = xy @ np.array([tilt, -0.5 * tilt]) + rng.normal(size=nrows) * 5.0
    uvw = np.column_stack([xy, w])
    freqs = np.linspace(130e6, 150e6, nchan) if nchan > 1 else np.array([140e6])
    # SVD_MFS.py:38-52
    u = uvw[:, 0:1] * freqs / C
    v = uvw[:, 1:2] * freqs / C
    ww = uvw[:, 2:3] * freqs / C
    D = np.column_stack([u.ravel(order="F"), v.ravel(order="F"), ww.ravel(order="F")])
    # SVD_MFS.py:60-75
    M1 = D.mean(axis=0)
    pca = PCA(2)
    pca.fit(D)
    V1 = pca.components_
    B1 = pca.transform(D)
    # SVD_MFS.py:98-110
    Vin = np.zeros((3, 3))
    Vin[:2] = V1
    Vin[2] = np.cross(Vin[0], Vin[1])
    Bin = B1.copy()
    Bin[:, 0] += M1 @ Vin[0]
    Bin[:, 1] += M1 @ Vin[1]
    return uvw, freqs, Bin, Vin


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for k, (nrows, nchan, tilt) in enumerate([(12, 1, 0.05), (2016, 1, 0.1), (500, 4, 0.2), (3000, 1, -0.3)]):
        uvw, freqs, B, V = case(rng, nrows, nchan, tilt)
        out[f"uvw{k}"] = uvw
        out[f"freqs{k}"] = freqs
        out[f"B{k}"] = B
        out[f"V{k}"] = V
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "pca_fixtures.npz"), **out)
    print("wrote", len(out) // 4, "cases")


if __name__ == "__main__":
    main()
