"""Write float32 golden vectors from the reference's own NumPy solver.

Run in the build container only (the reference is not present elsewhere):

    python tests/golden/make_golden.py

It imports ``shearwave`` from ``/root/reference/src`` and records, for small
seeded inputs, what the reference returns in its native float32 storage. These
pin the two things no closed form fixes: the reference's float32 rounding
(including NumPy 2 scalar promotion when ``dT`` is an ``np.float64``) and the
iteration at which its CG stops. Only inputs and outputs are written
(``ref_fp32.npz``); nothing of the reference travels with them.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from tests.golden.cases import CASES, build_case  # noqa: E402

REF_SRC = "/root/reference/src"


def main():
    sys.path.insert(0, REF_SRC)
    from shearwave import solver as ref  # noqa: PLC0415

    out = {}
    for name in CASES:
        c = build_case(name)
        if c["kind"] == "project":
            fx, fy, fz = ref.project_vector_field(
                c["fx"].copy(), c["fy"].copy(), c["fz"].copy(), *c["h"], c["tol"], c["max_iters"]
            )
            out[f"{name}/fx"], out[f"{name}/fy"], out[f"{name}/fz"] = fx, fy, fz
            continue
        snaps = []
        stride = c.get("stride", 0)

        def cb(it, u, snaps=snaps, stride=stride):
            if stride and (it + 1) % stride == 0:
                snaps.append(np.array(u, copy=True))

        u, v = ref.shear_fdtd(
            c["bx"], c["by"], c["bz"], c["rho"], c["mu"], *c["h"], c["dT"], c["n_steps"],
            opts=c["opts"], callback=cb,
        )
        out[f"{name}/u"], out[f"{name}/v"] = u, v
        if snaps:
            out[f"{name}/snaps"] = np.stack(snaps)
    path = os.path.join(HERE, "ref_fp32.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sum(a.nbytes for a in out.values()), "bytes uncompressed")


if __name__ == "__main__":
    main()
