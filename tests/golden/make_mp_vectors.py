"""Write tests/golden/mp_vectors.npz: arbitrary-precision Mie golden vectors.

Inputs are a fixed grid of (m, x) homogeneous and (m1, m2, x, y) coated
cases; outputs are the six efficiencies (mie_props.py:28-70 definitions)
and S1/S2 at five angles, from oracle/mpmie.py at high precision, both with
the reference's D_n start index (mie_coeffs.py:101, 204-205) and with a
converged start.  Run:  python tests/golden/make_mp_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import mie as M  # noqa: E402
from oracle import mpmie  # noqa: E402

HOMOG_M = [1.53 + 0.005j, 1.33 + 1e-6j, 1.6 + 0.01j, 1.8 + 1.0j, 1.45 + 0j]
HOMOG_X = [0.02, 0.5, 2.5, 12.0, 50.0, 209.0, 419.0, 1000.0, 2100.0]
COAT_M1 = 1.95 + 0.79j
COAT_M2 = [1.45 + 0j, 1.36 + 0.01j, 1.5 + 0.5j]
COAT_Y = [0.05, 1.0, 5.0, 30.0, 100.0, 209.0]
COAT_R = [0.3, 0.6, 0.9]
THETA = np.array([0.0, 30.0, 90.0, 150.0, 180.0])


def props_s12(an, bn, nmax, size):
    pr = M.mie_props_raw(an, bn, nmax, size)
    P, T = M.pt_table(np.cos(np.radians(THETA)), nmax)
    s1, s2 = M.s12_from_pt(an, bn, P, T)
    return np.array(pr), s1, s2


def main():
    rows = {"h_m": [], "h_x": [], "h_props_ref": [], "h_props_conv": [], "h_s1_ref": [],
            "h_s2_ref": [], "c_m1": [], "c_m2": [], "c_x": [], "c_y": [], "c_props_ref": [],
            "c_props_conv": [], "c_s1_ref": [], "c_s2_ref": []}
    for m in HOMOG_M:
        for x in HOMOG_X:
            an, bn, nm = mpmie.homog(m, x, start="reference")
            pr, s1, s2 = props_s12(an, bn, nm, x)
            an2, bn2, _ = mpmie.homog(m, x, start="converged")
            pr2 = np.array(M.mie_props_raw(an2, bn2, nm, x))
            rows["h_m"].append(m)
            rows["h_x"].append(x)
            rows["h_props_ref"].append(pr)
            rows["h_props_conv"].append(pr2)
            rows["h_s1_ref"].append(s1)
            rows["h_s2_ref"].append(s2)
            print("homog", m, x, nm, flush=True)
    for m2 in COAT_M2:
        for y in COAT_Y:
            for r in COAT_R:
                x = r * y
                an, bn, nm = mpmie.coated(COAT_M1, m2, x, y, start="reference")
                pr, s1, s2 = props_s12(an, bn, nm, y)
                an2, bn2, _ = mpmie.coated(COAT_M1, m2, x, y, start="converged")
                pr2 = np.array(M.mie_props_raw(an2, bn2, nm, y))
                rows["c_m1"].append(COAT_M1)
                rows["c_m2"].append(m2)
                rows["c_x"].append(x)
                rows["c_y"].append(y)
                rows["c_props_ref"].append(pr)
                rows["c_props_conv"].append(pr2)
                rows["c_s1_ref"].append(s1)
                rows["c_s2_ref"].append(s2)
                print("coated", m2, y, r, nm, flush=True)
    out = {k: np.array(v) for k, v in rows.items()}
    out["theta"] = THETA
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "mp_vectors.npz"),
                        **out)


if __name__ == "__main__":
    main()
