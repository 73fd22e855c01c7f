"""Trains the learned IAS15 start on the GPU and writes assist_amd/models/step_mlp.npz.

Workload (the uneven batch): 65,536 systems -- 70% main-belt, 30% NEO-like -- with
log-uniform spans of 30-10,000 days, half integrated backward, 10% with Marsden parameters,
on the seed-0 synthetic ephemeris.  The plain integrator records every accepted step after
the first 8 of each system (the start-up ramp carries no information about the settled step)
up to 10^7 records; features are computed on the GPU; the network is trained with Adam by the
library's own forward/backward kernels.

    python -m assist_amd.train_step_model [--records N] [--epochs E] [--out PATH]
"""
import argparse
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=65536)
    ap.add_argument("--records", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    import assist_amd
    from assist_amd import batch, learned as L, synthetic
    paths = synthetic.ensure_synthetic_ephemeris(0)
    ephem = assist_amd.Ephem(*paths)
    model = synthetic.SyntheticModel(0)
    rng = np.random.default_rng(args.seed)
    n = args.systems
    st, _ = synthetic.mixed_population(model, n, rng, 0.3, 0.0)
    states = st.reshape(n, 1, 6)
    t1 = synthetic.log_uniform_spans(n, rng)
    params = np.zeros((n, 1, 3))
    params[:, 0] = synthetic.marsden_params(n, rng, 0.1)
    cfg = batch.BatchConfig.default(min_dt=0.001, max_steps=200000)
    ts = time.perf_counter()
    rec, _, status, steps, evals = L.integrate_traced(ephem, np.zeros(n), t1, states, params, config=cfg,
                                                      capacity=args.records, every=1, skip=8)
    print("traces: %d records from %d systems (%d steps, %d evaluations, %d non-zero status) in %.1f s"
          % (rec.shape[0], n, steps.sum(), evals.sum(), (status != 0).sum(), time.perf_counter() - ts), flush=True)
    ts = time.perf_counter()
    X, _ = L.features(ephem, rec)
    T = L.targets(rec)
    keep = np.isfinite(X).all(axis=1) & np.isfinite(T).all(axis=1)
    X, T = X[keep], T[keep]
    print("features and targets: %d rows in %.1f s; log10 step mean %.3f sd %.3f"
          % (X.shape[0], time.perf_counter() - ts, T[:, 0].mean(), T[:, 0].std()), flush=True)
    # hold out 2% for validation
    nval = max(1, X.shape[0] // 50)
    Xv, Tv, Xt, Tt = X[:nval], T[:nval], X[nval:], T[nval:]
    net = L.StepModel(max_batch=max(args.batch, 65536), seed=args.seed)
    ts = time.perf_counter()
    hist = L.train(net, Xt, Tt, epochs=args.epochs, batch=args.batch, lr=args.lr, lr_end=args.lr / 20,
                   seed=args.seed, log=lambda s: print(s, flush=True))
    print("training: %.1f s" % (time.perf_counter() - ts), flush=True)
    d_Xv = batch.DeviceArray(Xv)
    d_Yv = batch.DeviceArray(np.zeros((nval, L.N_OUT), np.float32))
    for r0 in range(0, nval, 65536):
        net.forward(d_Xv, min(65536, nval - r0), d_Yv, row0=r0)
    Yv = d_Yv.download()
    err = Yv[:, 0] - Tv[:, 0]
    print("validation: log10 step error rms %.4f (step within x%.3f rms), b rms error %.4f (target rms %.4f)"
          % (np.sqrt(np.mean(err ** 2)), 10 ** np.sqrt(np.mean(err ** 2)),
             np.sqrt(np.mean((Yv[:, 1:22] - Tv[:, 1:22]) ** 2)), np.sqrt(np.mean(Tv[:, 1:22] ** 2))), flush=True)
    out = args.out or L.default_model_path()
    os.makedirs(os.path.dirname(out), exist_ok=True)
    net.save(out, loss_history=np.array(hist), records=rec.shape[0], systems=n, seed=args.seed)
    print("wrote", out, flush=True)


if __name__ == "__main__":
    sys.exit(main())
