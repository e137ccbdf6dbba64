"""Write the step golden vectors from PySIT's own step kernels.

Run in the container that holds the PySIT sources, after
``make -C oracle ref`` has compiled the reference templates
(pysit/solvers/constant_density_acoustic/time/scalar/
constant_density_acoustic_time_scalar_{2D,3D}.h) into ``oracle/_ref``:

    python tests/golden/make_golden.py

Output: ``tests/golden/step2d.npz`` and ``tests/golden/step3d.npz`` with
seeded random inputs (fields, wave speed, dense rhs, asymmetric PML arrays of
the reference's quadratic profile) and the reference's outputs for orders
2, 4, 6, 8 in float64 and float32.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def pml(n, amp, side):
    s = amp * np.linspace(0.0, 1.0, n + 1) ** 2
    return s[::-1][:-1].copy() if side == "left" else s[1:].copy()


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    rng = np.random.default_rng(20261202)
    out = {}
    nx, nz = 13, 11
    n = nx * nz
    ins = [rng.standard_normal(n) for _ in range(6)]
    ins[4] = 1.0 + 0.5 * rng.random(n)
    pm = [pml(3, 40.0, "left"), pml(4, 55.0, "right"), pml(2, 70.0, "left"), pml(3, 30.0, "right")]
    dt, dx, dz = 1.3e-3, 0.011, 0.013
    names = ["km1_u", "k_phix", "k_phiz", "k_u", "C", "rhs"]
    for nm, a in zip(names, ins):
        out[nm] = a
    for nm, a in zip(["xl", "xr", "zl", "zr"], pm):
        out[nm] = a
    out["scalars"] = np.array([dt, dx, dz, nx, nz], dtype=np.float64)
    for dtype, tag in ((np.float64, "f64"), (np.float32, "f32")):
        for order in (2, 4, 6, 8):
            res = O.step2d(order, *ins, *pm, dt, dx, dz, nx, nz, dtype=dtype, use_ref=True)
            for nm, a in zip(["kp1_phix", "kp1_phiz", "kp1_u"], res):
                out[f"{nm}_{tag}_o{order}"] = a.astype(dtype)
    np.savez_compressed(os.path.join(HERE, "step2d.npz"), **out)

    out = {}
    nx, ny, nz = 9, 10, 11
    n = nx * ny * nz
    ins = [rng.standard_normal(n) for _ in range(8)]
    ins[6] = 1.0 + 0.5 * rng.random(n)
    pm = [pml(2, 40.0, "left"), pml(3, 50.0, "right"), pml(3, 60.0, "left"), pml(2, 45.0, "right"),
          pml(2, 35.0, "left"), pml(3, 65.0, "right")]
    dt, dx, dy, dz = 1.1e-3, 0.011, 0.012, 0.013
    names = ["km1_u", "k_phix", "k_phiy", "k_phiz", "k_psi", "k_u", "C", "rhs"]
    for nm, a in zip(names, ins):
        out[nm] = a
    for nm, a in zip(["xl", "xr", "yl", "yr", "zl", "zr"], pm):
        out[nm] = a
    out["scalars"] = np.array([dt, dx, dy, dz, nx, ny, nz], dtype=np.float64)
    for dtype, tag in ((np.float64, "f64"), (np.float32, "f32")):
        for order in (2, 4, 6, 8):
            res = O.step3d(order, *ins, *pm, dt, dx, dy, dz, nx, ny, nz, dtype=dtype, use_ref=True)
            for nm, a in zip(["kp1_phix", "kp1_phiy", "kp1_phiz", "kp1_psi", "kp1_u"], res):
                out[f"{nm}_{tag}_o{order}"] = a.astype(dtype)
    np.savez_compressed(os.path.join(HERE, "step3d.npz"), **out)
    print("wrote", os.path.join(HERE, "step2d.npz"), os.path.join(HERE, "step3d.npz"))


if __name__ == "__main__":
    main()
