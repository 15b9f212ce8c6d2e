"""Write tests/golden/small_fp64.npz: seismograms and energy of two small runs of the
FP64 oracle (the CPU restatement of seismicCPML_3d_isotropic_simpelmodel.cpp:375-622),
kept as a regression fixture for the oracle and for the GPU FP64 path.

The reference holds no output files or test vectors of its own, so these are the
oracle's outputs on deterministic inputs (homogeneous model, no random numbers):

* "clean": the default table forms, Ricker force source in y, rigid faces;
* "reference": the reference's own choices (all compatibility quirks, the sigma
  source of CPU:620-622, the derivative-of-Gaussian wavelet of CPU:610, no rigid faces).

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from ela3d_amd import (COMPAT_REFERENCE, SRC_FORCE, SRC_REFERENCE_SIGMA, WAVELET_DGAUSSIAN,  # noqa: E402
                       WAVELET_RICKER, Config, homogeneous_model)

NX, NY, NZ, NSTEP = 29, 23, 26, 60


def receivers():
    return np.asarray([(i, NY // 2, NZ // 3) for i in range(0, NX, 3)] +
                      [(NX // 2, j, (2 * NZ) // 3) for j in range(0, NY, 4)], dtype=np.int32)


def configs():
    base = dict(nx=NX, ny=NY, nz=NZ, dx=10.0, dy=12.0, dz=9.0, dt=1.0e-3, nstep=NSTEP, npoints_pml=4,
                f0=18.0, factor=1e7, isrc=NX // 2, jsrc=NY // 2, ksrc=NZ // 2, receivers=receivers(),
                energy_every=1)
    return {
        "clean": Config(**base, src_kind=SRC_FORCE, wavelet=WAVELET_RICKER, compat_flags=0, dirichlet=1),
        "reference": Config(**base, src_kind=SRC_REFERENCE_SIGMA, wavelet=WAVELET_DGAUSSIAN,
                            compat_flags=COMPAT_REFERENCE, dirichlet=0),
    }


def model():
    return homogeneous_model(NX, NY, NZ, 3300.0, 3300.0 / 1.732, 2800.0)


def main():
    out, meta = {}, {}
    for name, cfg in configs().items():
        seis, energy, _ = oracle.run(cfg, model(), dtype=np.float64)
        out[f"{name}_seis"] = seis
        out[f"{name}_energy"] = energy
        meta[name] = {"src_kind": cfg.src_kind, "wavelet": cfg.wavelet, "compat_flags": cfg.compat_flags,
                      "dirichlet": cfg.dirichlet}
    meta["grid"] = {"nx": NX, "ny": NY, "nz": NZ, "nstep": NSTEP, "model": "homogeneous 3300, 3300/1.732, 2800"}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_fp64.npz")
    np.savez_compressed(path, meta=np.asarray(json.dumps(meta)), **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
