"""Write tests/golden/gsf_vectors.npz from the reference's own GSF program.

Runs oracle/_ref/spher_expan.x (compiled from src/gsf/spher_expan.f by
oracle/Makefile) on synthetic phase matrices in the 371-angle runoptics grid
(dointegration.py:739-742), written with np.savetxt exactly as the
reference's convertncdf.py does (:177-189), and stores the '.expan_coeff'
rows, the '.expan_matr' rows (the fitted matrix back at the input angles,
E15.5 / F11.5 text) and the fit error the program prints (ERREVAL, MAXABS).

Inputs: Rayleigh, three Henyey-Greenstein-like matrices, and four Mie
size-distribution matrices computed by the numpy oracle.
Run:  python tests/golden/make_gsf_vectors.py   (needs oracle/_ref built)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import lut, mie as M  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "spher_expan.x")


def run_ref(ang, f):
    """f: [6, N] in program order F11, F22, F33, F44, F12, F34.  Returns the
    '.expan_coeff' rows, CNORM, the '.expan_matr' rows (angle, F11..F34 back
    from the coefficients) and the fit error the program prints."""
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "m.txt")
        np.savetxt(fn, np.vstack([ang, f]).T)
        r = subprocess.run([EXE, fn], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                           check=True, text=True)
        rows = np.loadtxt(fn + ".expan_coeff", skiprows=1, unpack=True)
        head = open(fn + ".expan_coeff").readline().split()
        matr = np.loadtxt(fn + ".expan_matr", unpack=True)
    fit = [ln for ln in r.stdout.splitlines() if "FINAL NG" in ln][0].split()
    return rows[1:], float(head[1]), matr, float(fit[-1])


def matrices():
    ang = lut.angle_grid()
    mu = np.cos(np.radians(ang))
    out = []
    p11 = 0.75 * (1 + mu ** 2)
    out.append(("rayleigh", np.stack([p11, p11, 1.5 * mu, 1.5 * mu, -0.75 * (1 - mu ** 2), 0 * mu])))
    for g in (0.3, 0.7, 0.9):
        hg = (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5
        out.append(("hg%.1f" % g, np.stack([hg, 0.9 * hg, 0.8 * hg, 0.7 * hg, -0.1 * hg * (1 - mu ** 2),
                                            0.05 * hg * (1 - mu ** 2)])))
    # Mie lognormal size distributions (oracle), normalised as runoptics does
    for (rm, sg, lam, m) in [(0.05, 1.6, 0.55, 1.53 + 0.005j), (0.2, 2.0, 0.55, 1.45 + 0.01j),
                             (0.5, 2.0, 1.0, 1.53 + 0.005j), (0.1, 1.8, 0.35, 1.6 + 0.1j)]:
        xx = np.logspace(np.log10(0.01), np.log10(60.0), 160)
        dx = lut.getDR(xx)
        psd, _ = lut.lognorm_psd(rm, sg, xx, dx, lam, 1e9, 0.0)
        res = M.size_range(xx, mu, m.real, m.imag)
        raw = lut.raw_mie(xx, lam, res)
        integ = lut.integrate_psd(xx, [raw], [psd], [1.0], lam, [1e-6], 1000., 1000.)
        post = lut.post_integration(integ, ang)
        out.append(("mie_%g_%g" % (rm, sg), np.stack([post[k] for k in
                                                      ("p11", "p22", "p33", "p44", "p12", "p34")])))
    return ang, out


def main():
    ang, mats = matrices()
    names, F, A, C, MT, FE = [], [], [], [], [], []
    for name, f in mats:
        al, cn, mt, fe = run_ref(ang, f)
        names.append(name)
        F.append(f)
        A.append(al)
        C.append(cn)
        MT.append(mt)
        FE.append(fe)
        print(name, cn, al[0, :4], fe)
    np.savez_compressed(os.path.join(HERE, "gsf_vectors.npz"), names=np.array(names), ang=ang,
                        F=np.array(F), coeffs=np.array(A), cnorm_text=np.array(C),
                        matr=np.array(MT), fiterr=np.array(FE))


if __name__ == "__main__":
    main()
