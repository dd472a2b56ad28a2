"""Train the LUT surrogate on the GPU and save it.

    python -m geosmie_amd.train_surrogate --out geosmie_amd/data/surrogate_lut.npz

Builds the training tables with the GPU LUT sweep (geosmie_amd.lut_sweep) on
the full (m_r, m_i, lambda, RH, PSD) grid (or a strided subset), trains the
6-256-256-256-256-35 network with the HIP forward/backward/Adam kernels, fits
the trust gate on random validation slabs, and writes the model (parameters,
standardisation, gate) as an .npz file.  Progress goes to stdout.
"""
import argparse
import json
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--init", default=None,
                    help="continue training this model (its standardisation is kept)")
    ap.add_argument("--regate", default=None,
                    help="instead of training: load this model, re-fit its gate on fresh "
                         "validation slabs (--val-slabs) and write it to --out")
    ap.add_argument("--steps", type=int, default=400000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--stride", type=int, nargs=4, default=(1, 1, 1, 1),
                    help="subsampling of the m_r, m_i, lambda, gf axes")
    ap.add_argument("--val-slabs", type=int, default=30000)
    ap.add_argument("--safety", type=float, default=1.0,
                    help="a gate cell is trusted when every validation entry in it is within "
                         "bound / safety")
    ap.add_argument("--cells", type=int, nargs=6, default=None,
                    help="gate cells along ln r_mode, sigma_g, m_r, ln m_i, ln lambda, f_w")
    ap.add_argument("--cw-head", type=float, default=1000.0,
                    help="loss weight of Qext, SSA, g relative to each moment output")
    ap.add_argument("--aug-psd", type=int, default=40,
                    help="random (r_mode, sigma_g) pairs of a second pass over the grid's slabs")
    ap.add_argument("--weight-terms", type=int, default=2,
                    help="BF16 terms per network weight on the matrix cores (1 or 2)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="fp32",
                    help="S1/S2 precision of the training tables (Q and g are FP64 either way)")
    args = ap.parse_args(argv)
    from geosmie_amd import surrogate as S
    t0 = time.time()
    if args.regate:
        sur = S.Surrogate.load(args.regate)
        rep = S.regate(sur, n_val_slabs=args.val_slabs, seed=args.seed + 11, cells=args.cells,
                       safety=args.safety, log=lambda s: print(s, flush=True))
        sur.save(args.out)
        print(json.dumps(rep), flush=True)
        return
    sur, rep = S.train(train_stride=tuple(args.stride), n_val_slabs=args.val_slabs,
                       steps=args.steps, batch=args.batch, lr=args.lr, lr_final=args.lr_final,
                       cw_head=args.cw_head, seed=args.seed, precision=args.precision,
                       safety=args.safety,
                       n_aug_psd=args.aug_psd, weight_terms=args.weight_terms,
                       init=S.Surrogate.load(args.init) if args.init else None,
                       log=lambda s: print(s, flush=True))
    rep["total_s"] = time.time() - t0
    sur.info = rep
    sur.save(args.out)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
