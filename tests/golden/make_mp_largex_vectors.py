"""Write tests/golden/mp_largex_vectors.npz: arbitrary-precision Mie vectors
for the large-size-parameter angular sums and the coated start correction.

1. Homogeneous spheres at size parameters 250..2094 (series lengths up to
   2148) with refractive indices spanning the large-x workload's range
   (m_r 1.33-1.60, m_i 1e-6-1e-2): the six efficiencies and S1/S2 at 12
   scattering angles from forward through side to exact back scattering,
   with a_n, b_n from oracle/mpmie.py (the reference's D_n start,
   mie_coeffs.py:101) and the angular sums carried out in mpmath
   (mpmie.s12), so no float64 dot product enters the expected values.
2. Coated spheres whose shell is optically denser than the core
   (|m2| >= |m1|: m1 = 1.33, 1.33+1e-4i; m2 = 1.5, 1.5+1e-3i, 1.6+5e-3i),
   y = 6..209 and core/shell ratios 0.3, 0.7, 0.95.  There the reference's
   common start index nmx = round(max(n_max, |m1 y|, |m2 y|) + 16)
   (mie_coeffs.py:203-205) lies within a few Airy widths of |m2 x| or
   |m2 y|, so its D_n(m2 x), D_n(m2 y) are not the converged values; the
   expected values reproduce that start exactly (start="reference") and
   the converged ones are stored beside them.

Run:  python tests/golden/make_mp_largex_vectors.py   (about a minute)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import mie as M  # noqa: E402
from oracle import mpmie  # noqa: E402

THETA = np.array([0.0, 0.5, 2.0, 10.0, 45.0, 90.0, 120.0, 150.0, 170.0, 178.0, 179.5, 180.0])
HX_X = [250.0, 600.0, 1000.0, 1500.0, 2094.0]
HX_M = [1.33 + 1e-6j, 1.45 + 1e-4j, 1.60 + 1e-2j, 1.38 + 3e-3j]
CD_M1 = [1.33 + 0j, 1.33 + 1e-4j]
CD_M2 = [1.5 + 0j, 1.5 + 1e-3j, 1.6 + 5e-3j]
CD_Y = [6.0, 12.0, 30.0, 60.0, 100.0, 150.0, 209.0]
CD_R = [0.3, 0.7, 0.95]


def main():
    cost = np.cos(np.radians(THETA))
    out = {k: [] for k in ("hx_x", "hx_m", "hx_nmax", "hx_props", "hx_s1", "hx_s2",
                           "cd_m1", "cd_m2", "cd_x", "cd_y", "cd_props", "cd_props_conv",
                           "cd_s1", "cd_s2")}
    for x in HX_X:
        for m in HX_M:
            an, bn, nm = mpmie.homog(m, x, start="reference")
            s1, s2 = mpmie.s12(an, bn, cost)
            out["hx_x"].append(x)
            out["hx_m"].append(m)
            out["hx_nmax"].append(nm)
            out["hx_props"].append(M.mie_props_raw(an, bn, nm, x))
            out["hx_s1"].append(s1)
            out["hx_s2"].append(s2)
            print("homog", x, m, nm, flush=True)
    for m1 in CD_M1:
        for m2 in CD_M2:
            for y in CD_Y:
                for r in CD_R:
                    x = r * y
                    an, bn, nm = mpmie.coated(m1, m2, x, y, start="reference")
                    s1, s2 = mpmie.s12(an, bn, cost)
                    an2, bn2, _ = mpmie.coated(m1, m2, x, y, start="converged")
                    out["cd_m1"].append(m1)
                    out["cd_m2"].append(m2)
                    out["cd_x"].append(x)
                    out["cd_y"].append(y)
                    out["cd_props"].append(M.mie_props_raw(an, bn, nm, y))
                    out["cd_props_conv"].append(M.mie_props_raw(an2, bn2, nm, y))
                    out["cd_s1"].append(s1)
                    out["cd_s2"].append(s2)
                    print("coated", m1, m2, y, r, nm, flush=True)
    arr = {k: np.array(v) for k, v in out.items()}
    arr["theta"] = THETA
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                     "mp_largex_vectors.npz"), **arr)


if __name__ == "__main__":
    main()
