"""Train the learned Davidson preconditioner and measure it against the
diagonal one.

Data: for each training seed, a matrix of the class (dense, default: the
N=16384 class d ~ U[0,50] sorted with symmetric N(0,1) * 0.1/sqrt(N)
off-diagonal; or, with --matrix sparse, the orbital-sum CSR model), built by
the library's device generators, its exact lowest pairs from the library's
block solver, and the Ritz states of the library's own solver: the same
solve (one block IPT step, then Davidson) stopped after s = 0, 1, 2, ...
Davidson steps, each returning its pairs, until the tolerance or the step
cap (--states library, the default).  --states torch instead re-expresses
the Davidson iteration with torch on the device (the library's restart rule
and residual scaling) and records every step of one run.  At every state
and for every unconverged target j the
pairs are (features of component i, q_i = r_i / s_i) -> t_i, the component
of the Jacobi-Davidson correction u*_j / (u*_j . x_j) - x_j after projection
off the current Ritz vectors; the loss weights each target column by
1 / ||t||^2.  Components near theta_j (the 64 nearest diagonal entries on
each side) are all kept; the rest are sampled at 1/32 with weight 32.

Training runs on the library's HIP gradient and Adam kernels.  A second
round collects states from Davidson runs that use the current network
(guarded by the diagonal correction, as in the library) and trains on the
union.  The model is written in the library's model-file format and then
measured with the library's own block solver on the test seeds:
Davidson steps and single-column products to reach the residual
tolerance, learned vs diagonal.

    python -m block_ipt_amd.mlp_train --out block_ipt_amd/davidson_mlp.bin
    python -m block_ipt_amd.mlp_train --matrix sparse --n 200000 --k 32 --test-n 1000000 \
        --out block_ipt_amd/davidson_mlp_sparse.bin

The script needs a GPU; torch provides device memory and the dense algebra
of the data generation only (sparse products go through the library's CSR
product).
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from block_ipt_amd import api

P = api.MLP_NPARAM


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sync():
    torch.cuda.synchronize()


def ptr(t):
    return t.data_ptr()


def _reference_env():
    # reference pairs: block solve with Davidson refinement to 1e-12
    os.environ.update({"IPT_BLOCK_CLUSTER": "1", "IPT_DAVIDSON_ENRICH": "1", "IPT_DAVIDSON_STEPS": "200",
                       "IPT_DAVIDSON_CONVERGED_TOL": "1e-12"})
    for v in ("IPT_PRECISION", "IPT_BLOCK_CLUSTER_MAXITER", "IPT_DAVIDSON_PRECOND"):
        os.environ.pop(v, None)


def _clear_reference_env():
    for v in ("IPT_DAVIDSON_ENRICH", "IPT_DAVIDSON_STEPS", "IPT_DAVIDSON_CONVERGED_TOL"):
        os.environ.pop(v, None)


class DenseMatrix:
    """Dense class: diag(sorted U[0,50]) + symmetric N(0,1) vscale/sqrt(N)."""

    def __init__(self, n, seed, vscale):
        self.n = n
        self.H = torch.empty((n, n), dtype=torch.float64, device="cuda")  # column-major == H^T == H
        sync()
        api.generate_dense(ptr(self.H), n, seed, 0.0, 50.0, vscale)
        self.d = torch.diagonal(self.H).clone()
        self.s = self.H.abs().sum(0) - self.d.abs()
        self.normH = float(self.H.abs().sum(0).max())

    def apply(self, X):
        return self.H @ X

    def solve(self, k, maxiter):
        """library block solve on the device: (vectors n x k column-major, values)"""
        sync()
        dv, dl, _, _ = api.ipt_cuda_device_tol(ptr(self.H), self.n, k, 1e-12, maxiter)
        return dv, dl


class SparseMatrix:
    """Sparse orbital-sum class (40 orbital energies U[-20,5], diagonal = sums
    of 8, m drawn columns per row with N(0, 0.02^2) values, mirrored)."""

    def __init__(self, n, seed, m):
        self.n = n
        cap = n * (2 * m + 1)
        self.rp = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        ci = torch.empty(cap, dtype=torch.int32, device="cuda")
        vv = torch.empty(cap, dtype=torch.float64, device="cuda")
        sync()
        self.nnz = api.generate_orbital_csr(n, m, seed, ptr(self.rp), ptr(ci), ptr(vv), cap)
        self.ci = ci[:self.nnz].clone()
        self.v = vv[:self.nnz].clone()
        del ci, vv
        rows = torch.repeat_interleave(torch.arange(n, device="cuda"), (self.rp[1:] - self.rp[:-1]).long())
        absum = torch.zeros(n, dtype=torch.float64, device="cuda").index_add_(0, rows, self.v.abs())
        self.d = torch.zeros(n, dtype=torch.float64, device="cuda")
        diag = rows == self.ci.long()
        self.d[rows[diag]] = self.v[diag]
        self.s = absum - self.d.abs()
        self.normH = float(absum.max())

    def apply(self, X):
        X = X.contiguous()  # row-major n x c: the layout of the CSR product
        Y = torch.empty_like(X)
        sync()
        api.csr_spmm(ptr(self.rp), ptr(self.ci), ptr(self.v), self.n, ptr(X), X.shape[1], ptr(Y), X.shape[1],
                     X.shape[1])
        return Y

    def solve(self, k, maxiter):
        sync()
        dv, dl, _, _ = api.ipt_cuda_sparse_csc_device_tol(ptr(self.rp), ptr(self.ci), ptr(self.v), self.n, k,
                                                          self.nnz, 1e-12, maxiter)
        return dv, dl


def make_matrix(args, seed):
    if args.matrix == "sparse":
        return SparseMatrix(args.n, seed, args.draws)
    return DenseMatrix(args.n, seed, args.vscale)


class Problem:
    """One matrix of the class with its exact lowest pairs."""

    def __init__(self, args, k, seed, extra=16):
        self.mat = make_matrix(args, seed)
        self.n, self.k = self.mat.n, k
        self.d, self.s, self.normH = self.mat.d, self.mat.s, self.mat.normH
        _reference_env()
        ke = k + extra
        dv, dl = self.mat.solve(ke, 400)
        hip = api.hip()
        U = hip.to_host(dv, (self.n, ke))
        lam = hip.to_host(dl, (ke,))
        api.free_device_result(dv, dl)
        _clear_reference_env()
        self.U = torch.tensor(U, device="cuda")
        # the reference pairs used for matching must be eigenpairs (scaled
        # residual); the last few of the k + extra may be less accurate
        lt = torch.tensor(lam, device="cuda")
        res = torch.linalg.norm(self.mat.apply(self.U) - self.U * lt, dim=0) / max(1.0, self.normH)
        self.ok = bool(res[:k + extra // 2].max() <= 1e-9)

    def apply(self, X):
        return self.mat.apply(X)


def rayleigh_ritz(V, HV, k):
    G = V.T @ HV
    G = 0.5 * (G + G.T)
    w, C = torch.linalg.eigh(G.cpu())
    C = C[:, :k].to(V.device)
    return w[:k].to(V.device), V @ C, HV @ C


def orth_against(V, T):
    """Two classical Gram-Schmidt passes against V, then an orthonormal basis
    of what is left (columns with a negligible remainder are dropped)."""
    for _ in range(2):
        T = T - V @ (V.T @ T)
    nrm0 = torch.linalg.norm(T, dim=0)
    Q, R = torch.linalg.qr(T)
    dg = torch.abs(torch.diagonal(R))
    keep = dg > 1e-10 * torch.clamp(nrm0.max(), min=1e-300)
    return Q[:, keep]


def diag_correction(R, d, th, clip):
    den = d[:, None] - th[None, :]
    den = torch.where(den.abs() < clip, torch.where(den < 0, -clip, clip), den)
    return R / den


class Net:
    """Device parameters plus the library kernels that use them."""

    def __init__(self, params):
        self.p = torch.tensor(params, dtype=torch.float32, device="cuda")
        self.m = torch.zeros(P, dtype=torch.float32, device="cuda")
        self.v = torch.zeros(P, dtype=torch.float32, device="cuda")
        self.g = torch.zeros(P, dtype=torch.float32, device="cuda")
        self.loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.step = 0

    def features(self, R, d, s, th):
        """features (nt, n, 3) and q (nt, n) of the columns of R."""
        n, nt = R.shape
        Rc = R.T.contiguous()  # column j of R at Rc[j] (column-major n x nt)
        idx = torch.arange(nt, dtype=torch.int32, device="cuda")
        feat = torch.empty((nt, n, 3), dtype=torch.float32, device="cuda")
        q = torch.empty((nt, n), dtype=torch.float32, device="cuda")
        thc = th.contiguous()
        sync()
        api.mlp_features(ptr(Rc), n, ptr(d), ptr(s), ptr(thc), ptr(idx), n, nt, ptr(feat), ptr(q))
        return feat, q

    def correction(self, R, d, s, th):
        feat, q = self.features(R, d, s, th)
        out = torch.empty_like(q)
        api.mlp_infer(ptr(self.p), ptr(feat), ptr(q), q.numel(), ptr(out))
        return out.T.double()

    def train(self, F, Q, T, W, steps, batch, lr0, seed):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        N = F.shape[0]
        hist = []
        for it in range(steps):
            idx = torch.randint(0, N, (batch,), device="cuda", generator=g)
            f, q, t, w = F[idx].contiguous(), Q[idx].contiguous(), T[idx].contiguous(), W[idx].contiguous()
            sync()
            api.mlp_gradient(ptr(self.p), ptr(f), ptr(q), ptr(t), ptr(w), batch, ptr(self.g), ptr(self.loss))
            self.step += 1
            lr = lr0 * 0.5 * (1.0 + np.cos(np.pi * it / steps))
            api.mlp_adam_step(ptr(self.p), ptr(self.m), ptr(self.v), ptr(self.g), 1.0 / batch, lr, self.step)
            if it % 500 == 0 or it == steps - 1:
                hist.append(float(self.loss.item()) / batch)
                log("    step %5d  loss/sample %.4e  lr %.2e" % (it, hist[-1], lr))
        return hist


def davidson_states(pb, k, net, conv_tol, max_steps):
    """Davidson from the identity start, the library's rules; returns the
    recorded (R, theta, X, cand) states and the step / trial counts."""
    n = pb.n
    d, s = pb.d, pb.s
    scale = max(1.0, pb.normH)
    X0 = torch.zeros((n, k), dtype=torch.float64, device="cuda")
    X0[torch.arange(k), torch.arange(k)] = 1.0
    V = X0
    HV = pb.apply(V)
    th, X, HX = rayleigh_ritz(V, HV, k)
    kmax = min(n, k + max(16, (2 * k + 2) // 3))
    retain = max(0, kmax - k - 1)
    clip = np.finfo(float).eps ** (2.0 / 3.0) * max(scale, float(th.abs().max()))
    recent = None
    hold = 0
    states, trials, steps = [], 0, 0

    def resid(X, HX, th):
        R = HX - X * th
        return R, torch.linalg.norm(R, dim=0) / (scale * torch.clamp(torch.linalg.norm(X, dim=0), min=1.0))

    R, res = resid(X, HX, th)
    for step in range(max_steps):
        cand = torch.nonzero(res > conv_tol).flatten()
        if cand.numel() == 0:
            break
        if V.shape[1] > k and V.shape[1] + cand.numel() > kmax:
            V, HV = X.clone(), HX.clone()
            if recent is not None and retain > 0:
                Q = orth_against(V, recent[:, :retain])
                if Q.shape[1] > 0:
                    V = torch.cat([V, Q], 1)
                    HV = torch.cat([HV, pb.apply(Q)], 1)
        lim = min(k, max(0, kmax - V.shape[1]))
        cand = cand[:lim]
        if cand.numel() == 0:
            break
        states.append((R[:, cand].clone(), th[cand].clone(), X[:, cand].clone(), X.clone()))
        before = float(res.max())
        accepted = None
        use_net = net is not None and hold == 0
        hold = max(0, hold - 1)
        for learned in ([True, False] if use_net else [False]):
            if learned:
                T = net.correction(R[:, cand], d, s, th[cand])
            else:
                T = diag_correction(R[:, cand], d, th[cand], clip)
            Q = orth_against(V, T)
            trials += 1
            if Q.shape[1] == 0:
                continue
            V2 = torch.cat([V, Q], 1)
            HV2 = torch.cat([HV, pb.apply(Q)], 1)
            th2, X2, HX2 = rayleigh_ritz(V2, HV2, k)
            R2, res2 = resid(X2, HX2, th2)
            if not learned or float(res2.max()) < before:
                accepted = (Q, V2, HV2, th2, X2, HX2, R2, res2)
                break
            hold = 2  # the library's rule: two diagonal steps after a rejection
        if accepted is None:
            break
        recent, V, HV, th, X, HX, R, res = accepted
        steps += 1
        if V.shape[1] >= kmax:
            V, HV = X.clone(), HX.clone()
            Q = orth_against(V, recent[:, :retain]) if retain > 0 else recent[:, :0]
            if Q.shape[1] > 0:
                V = torch.cat([V, Q], 1)
                HV = torch.cat([HV, pb.apply(Q)], 1)
    return states, steps, trials, float(res.max())


def library_states(args, pb, model, conv_tol, max_steps):
    """Ritz states of the library's solver: one block IPT step, then
    Davidson stopped after s = 0, 1, ... steps (diagonal corrections, or the
    guarded learned ones with `model`); the pairs each run returns (its best
    state so far) are recorded with their residuals until every target
    reaches conv_tol or s reaches max_steps.  Returns (states, steps of the
    last run, final largest residual)."""
    hip = api.hip()
    n, k = pb.n, pb.k
    scale = max(1.0, pb.normH)
    env = {"IPT_BLOCK_CLUSTER": "1", "IPT_BLOCK_CLUSTER_MAXITER": str(args.ipt_steps), "IPT_DAVIDSON_ENRICH": "1",
           "IPT_DAVIDSON_CONVERGED_TOL": repr(conv_tol)}
    for kk in ("IPT_PRECISION", "IPT_DAVIDSON_PRECOND", "IPT_DAVIDSON_MLP_MODEL"):
        os.environ.pop(kk, None)
    os.environ.update(env)
    if model is not None:
        os.environ["IPT_DAVIDSON_PRECOND"] = "mlp"
        os.environ["IPT_DAVIDSON_MLP_MODEL"] = model
    states, last, steps = [], None, 0
    try:
        for s_lim in range(0, max_steps + 1):
            os.environ["IPT_DAVIDSON_STEPS"] = str(s_lim)
            dv, dl = pb.mat.solve(k, args.ipt_steps)
            X = torch.tensor(hip.to_host(dv, (n, k)), device="cuda")
            th = torch.tensor(hip.to_host(dl, (k,)), device="cuda")
            api.free_device_result(dv, dl)
            steps = api.last_davidson_stats()["accepted"]
            R = pb.apply(X) - X * th
            res = torch.linalg.norm(R, dim=0) / (scale * torch.clamp(torch.linalg.norm(X, dim=0), min=1.0))
            last = float(res.max())
            cand = torch.nonzero(res > conv_tol).flatten()
            if cand.numel() == 0:
                break
            if steps < s_lim:
                break  # the run stopped before its limit (no further progress)
            states.append((R[:, cand].clone(), th[cand].clone(), X[:, cand].clone(), X.clone()))
    finally:
        for kk in ("IPT_DAVIDSON_PRECOND", "IPT_DAVIDSON_MLP_MODEL", "IPT_DAVIDSON_STEPS",
                   "IPT_BLOCK_CLUSTER_MAXITER", "IPT_DAVIDSON_CONVERGED_TOL", "IPT_DAVIDSON_ENRICH"):
            os.environ.pop(kk, None)
    return states, steps, last


def samples_from_states(pb, states, net_for_features, near=64, rate=32, seed=0):
    """Stratified (features, q, target, weight) samples of the states."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    n = pb.n
    out = [[], [], [], []]
    for (Rc, thc, Xc, Xall) in states:
        feat, q = net_for_features.features(Rc, pb.d, pb.s, thc)
        ov = pb.U.T @ Xc                       # (ke, nt)
        m = torch.argmax(ov.abs(), dim=0)
        us = pb.U[:, m] / ov[m, torch.arange(m.numel(), device="cuda")]
        e = us - Xc
        e = e - Xall @ (Xall.T @ e)            # JD-orthogonal correction
        nrm2 = (e * e).sum(0)
        pos = torch.searchsorted(pb.d, thc)
        ii = torch.arange(n, device="cuda")[None, :]
        nearm = (ii - pos[:, None]).abs() <= near
        pick = nearm | (torch.rand((thc.numel(), n), device="cuda", generator=gen) < 1.0 / rate)
        w = torch.where(nearm, 1.0, float(rate)) / nrm2[:, None]
        sel = pick & (nrm2[:, None] > 0)
        out[0].append(feat[sel])
        out[1].append(q[sel])
        out[2].append(e.T[sel].float())
        out[3].append(w[sel].float())
    return [torch.cat(o) for o in out]


def init_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    p[0:192] = rng.standard_normal(192) * np.sqrt(1.0 / 3.0)
    p[256:4352] = rng.standard_normal(4096) * np.sqrt(1.0 / 64.0)
    p[4416:4480] = rng.standard_normal(64) * np.sqrt(1.0 / 64.0)
    return p


def measure(args, model, seeds):
    """Library block solver, Davidson-dominant configuration (one block IPT
    step, then Davidson to the tolerance), diagonal vs learned corrections."""
    rows = []
    for seed in seeds:
        mat = make_matrix(args, seed)
        rec = {"seed": seed}
        for mode in ("diag", "mlp"):
            env = {"IPT_BLOCK_CLUSTER": "1", "IPT_BLOCK_CLUSTER_MAXITER": str(args.ipt_steps),
                   "IPT_DAVIDSON_ENRICH": "1", "IPT_DAVIDSON_STEPS": str(args.max_steps),
                   "IPT_DAVIDSON_CONVERGED_TOL": repr(args.conv_tol)}
            for kk in ("IPT_PRECISION", "IPT_DAVIDSON_PRECOND", "IPT_DAVIDSON_MLP_MODEL"):
                os.environ.pop(kk, None)
            os.environ.update(env)
            if mode == "mlp":
                os.environ["IPT_DAVIDSON_PRECOND"] = "mlp"
                os.environ["IPT_DAVIDSON_MLP_MODEL"] = model
            t0 = time.time()
            dv, dl = mat.solve(args.k, args.ipt_steps)
            el = time.time() - t0
            api.free_device_result(dv, dl)
            st = api.last_davidson_stats()
            rec[mode] = {"steps": st["accepted"], "trials": st["accepted"] + st["mlp_rejected"],
                         "matvecs": st["matvecs"], "residual": st["residual_after"], "time_s": el,
                         "mlp_rejected": st["mlp_rejected"]}
        for kk in ("IPT_DAVIDSON_PRECOND", "IPT_DAVIDSON_MLP_MODEL"):
            os.environ.pop(kk, None)
        log("  seed %d  diag %s  mlp %s" % (seed, rec["diag"], rec["mlp"]))
        rows.append(rec)
        del mat
    tot = {m: {f: sum(r[m][f] for r in rows) for f in ("steps", "trials", "matvecs")} for m in ("diag", "mlp")}
    return rows, tot


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "davidson_mlp.bin"))
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--vscale", type=float, default=0.1)
    ap.add_argument("--matrix", default="dense", choices=["dense", "sparse"],
                    help="dense: d ~ U[0,50] + N(0,1) vscale/sqrt(N); sparse: the orbital-sum CSR model")
    ap.add_argument("--draws", type=int, default=200, help="sparse: drawn columns per row")
    ap.add_argument("--test-n", type=int, default=0, help="matrix size of the measurement (default: --n)")
    ap.add_argument("--train-seeds", default="1-20")
    ap.add_argument("--test-seeds", default="100-109")
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=1 << 17)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--conv-tol", type=float, default=1e-10)
    ap.add_argument("--max-steps", type=int, default=120)
    ap.add_argument("--ipt-steps", type=int, default=1)
    ap.add_argument("--report", default=None, help="JSON file for the measurement")
    ap.add_argument("--measure-only", action="store_true")
    ap.add_argument("--states", default="library", choices=["library", "torch"],
                    help="training states: the library solver's (default) or a torch re-expression of Davidson")
    args = ap.parse_args(argv)

    def seeds(s):
        a, b = s.split("-")
        return list(range(int(a), int(b) + 1))

    torch.manual_seed(0)
    if not args.measure_only:
        net = Net(init_params(0))
        data = None
        for rnd in range(args.rounds):
            t0 = time.time()
            parts = []
            for sd in seeds(args.train_seeds):
                try:
                    pb = Problem(args, args.k, sd)
                except api.IptError as e:
                    log("  round %d seed %d: reference solve failed (%s), seed skipped" % (rnd, sd, e))
                    continue
                if not pb.ok:
                    log("  round %d seed %d: reference pairs not converged, seed skipped" % (rnd, sd))
                    continue
                if args.states == "library":
                    model = None
                    if rnd > 0:
                        model = args.out + ".round%d" % rnd
                        api.save_mlp_model(model, net.p.cpu().numpy())
                    states, steps, r = library_states(args, pb, model, args.conv_tol, args.max_steps)
                    trials = steps
                else:
                    states, steps, trials, r = davidson_states(pb, args.k, net if rnd > 0 else None, args.conv_tol,
                                                               args.max_steps)
                parts.append(samples_from_states(pb, states, net, seed=sd * 7 + rnd))
                log("  round %d seed %d: %d states, %d steps, %d trials, final residual %.2e"
                    % (rnd, sd, len(states), steps, trials, r))
                del pb
            new = [torch.cat([p[i] for p in parts]) for i in range(4)]
            data = new if data is None else [torch.cat([a, b]) for a, b in zip(data, new)]
            log("round %d: %d samples (%.1f s)" % (rnd, data[0].shape[0], time.time() - t0))
            net.train(data[0].contiguous(), data[1].contiguous(), data[2].contiguous(), data[3].contiguous(),
                      args.train_steps, args.batch, args.lr, seed=rnd)
        api.save_mlp_model(args.out, net.p.cpu().numpy())
        log("model written to %s" % args.out)
    if args.test_n > 0:
        args.n = args.test_n
    rows, tot = measure(args, args.out, seeds(args.test_seeds))
    red_steps = 1.0 - tot["mlp"]["steps"] / max(1, tot["diag"]["steps"])
    red_trials = 1.0 - tot["mlp"]["trials"] / max(1, tot["diag"]["trials"])
    red_mv = 1.0 - tot["mlp"]["matvecs"] / max(1, tot["diag"]["matvecs"])
    log("TOTAL diag %s mlp %s  fewer steps %.1f%%  fewer trials %.1f%%  fewer column products %.1f%%"
        % (tot["diag"], tot["mlp"], 100 * red_steps, 100 * red_trials, 100 * red_mv))
    if args.report:
        with open(args.report, "w") as f:
            json.dump({"matrix": args.matrix, "n": args.n, "k": args.k, "conv_tol": args.conv_tol,
                       "ipt_steps": args.ipt_steps,
                       "rows": rows, "total": tot, "fewer_steps": red_steps, "fewer_trials": red_trials,
                       "fewer_column_products": red_mv}, f, indent=1)


if __name__ == "__main__":
    main()
