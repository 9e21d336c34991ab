"""Learned warm start for a family of LPs that share one sparsity structure.

Extension with no counterpart in the reference (its PDLP always starts at
x = proj(0), y = 0; src/lp/first_order/pdlp.cpp:186-191).  Pipeline:

1. features: the instance's data perturbation relative to the family's base
   LP, d = [(b - b0) / (1 + |b0|), (c - c0) / (1 + |c0|), (u - u0) / (1 + |u0|)],
   projected on the top principal directions of the training perturbations
   (256 "summary features" of b, c and the bounds);
2. targets: the GPU PDLP solutions of the training instances, z = [x / u; y / sy]
   (the primal as a fraction of each variable's upper bound, so a variable at
   a bound maps to exactly 0 or 1 whatever the instance's bound; the dual over
   one global scale), projected on their top principal directions (64
   coefficients of a reduced solution basis);
3. a BF16 MLP on MFMA (mcpdlp.MLP: hand-written forward/backward/Adam kernels)
   maps standardized features to standardized coefficients;
4. predict(): reconstruct z from the predicted coefficients, hand (x0, y0) to
   solve_pdlp as warm_primal / warm_dual;
5. the start primal weight: the training solves' final weight (median), since
   a start near the optimum with a mismatched weight first drifts away from
   it, or the cold-start default -- tune_omega() picks the one that needs
   fewer iterations on a held-out validation batch;
6. solve_batch(): the warm-started batch with the cold start as the fallback.
   An instance that has not reached the tolerance within twice the iterations
   the hardest training solve needed is solved again from the cold start, so
   a bad prediction costs iterations, never convergence; the KKT score alone
   decides termination.
"""
from __future__ import annotations

import numpy as np

import mcpdlp as mp


def _rhs(lp: mp.LPData) -> np.ndarray:
    """One finite right-hand side per row (upper if finite, else lower)."""
    up = lp.row_upper["kind"] == mp.FINITE
    return np.where(up, lp.row_upper["value"], lp.row_lower["value"]).astype(np.float64)


class WarmStart:
    def __init__(self, base: mp.LPData, n_features: int = 256, n_coeffs: int = 64, hidden: int = 256,
                 layers: int = 3, seed: int = 1, device: int = 0):
        self.base = base
        self.n_features, self.n_coeffs = n_features, n_coeffs
        self.hidden, self.layers, self.seed, self.device = hidden, layers, seed, device
        self.b0, self.c0, self.u0 = _rhs(base), base.objective.astype(np.float64), base.var_upper["value"].copy()
        self.net = None

    def perturbation(self, lp: mp.LPData) -> np.ndarray:
        return np.concatenate([(_rhs(lp) - self.b0) / (1.0 + np.abs(self.b0)),
                               (lp.objective - self.c0) / (1.0 + np.abs(self.c0)),
                               (lp.var_upper["value"] - self.u0) / (1.0 + np.abs(self.u0))])

    def features(self, lps) -> np.ndarray:
        D = np.stack([self.perturbation(lp) for lp in lps])
        F = (D - self.d_mean) @ self.d_basis
        return ((F - self.f_mean) / self.f_std).astype(np.float32)

    @staticmethod
    def _pca(X: np.ndarray, k: int):
        """Mean and the top-k principal directions of the rows of X.  Only
        directions with a singular value above 1e-9 of the largest are kept:
        with fewer samples than components the rest span rounding noise, on
        which the training rows project to ~0 but a new row does not, and a
        standardized feature on such a direction would be O(1e10).  Missing
        directions are zero columns, which map every row to exactly 0."""
        mean = X.mean(axis=0)
        _, sv, vt = np.linalg.svd(X - mean, full_matrices=False)
        rank = int(np.sum(sv > 1e-9 * sv[0])) if sv.size and sv[0] > 0 else 0
        basis = vt[:min(k, rank)].T
        if basis.shape[1] < k:
            basis = np.pad(basis, ((0, 0), (0, k - basis.shape[1])))
        return mean, basis

    @staticmethod
    def _upper(lp: mp.LPData) -> np.ndarray:
        u = lp.var_upper["value"].astype(np.float64)
        return np.where((lp.var_upper["kind"] == mp.FINITE) & (u > 0), u, 1.0)

    def fit(self, lps, xs, ys, omegas=None, iterations=None, epochs: int = 300, batch: int = 64,
            lr: float = 1e-3) -> dict:
        D = np.stack([self.perturbation(lp) for lp in lps])
        self.d_mean, self.d_basis = self._pca(D, self.n_features)
        F = (D - self.d_mean) @ self.d_basis
        self.f_mean, self.f_std = F.mean(axis=0), F.std(axis=0) + 1e-12
        Fs = ((F - self.f_mean) / self.f_std).astype(np.float32)
        X = np.stack([x / self._upper(lp) for x, lp in zip(xs, lps)])
        Y = np.stack(ys)
        self.sy = float(np.sqrt(np.mean(Y * Y))) + 1e-12
        self.omega = float(np.median(omegas)) if omegas is not None and len(omegas) else 0.0
        # warm-start iteration cap before the cold fallback (0: no cap)
        self.cap = int(2 * max(iterations)) if iterations is not None and len(iterations) else 0
        Z = np.concatenate([X, Y / self.sy], axis=1)
        self.z_mean, self.z_basis = self._pca(Z, self.n_coeffs)
        Cf = (Z - self.z_mean) @ self.z_basis
        self.c_mean, self.c_std = Cf.mean(axis=0), Cf.std(axis=0) + 1e-12
        Cs = ((Cf - self.c_mean) / self.c_std).astype(np.float32)
        explained = 1.0 - np.sum((Z - self.z_mean - Cf @ self.z_basis.T) ** 2) / np.sum((Z - self.z_mean) ** 2)
        # mcp_mlp_train skips the tail that does not fill a mini-batch: with fewer
        # samples than a batch, one batch of all of them
        batch = max(1, min(batch, len(lps)))
        self.net = mp.MLP(self.n_features, self.hidden, self.layers, self.n_coeffs,
                          capacity=max(32, batch), seed=self.seed, device=self.device)
        loss, ms = self.net.train(Fs, Cs, batch=batch, epochs=epochs, lr=lr, seed=self.seed + 1)
        return {"train_loss": loss, "train_ms": ms, "pca_explained": float(explained),
                "samples": len(lps), "params": self.net.nparam, "primal_weight": self.omega,
                "iteration_cap": self.cap}

    def predict(self, lps, use_network: bool = True):
        """[(x0, y0)] per instance; pass self.omega as initial_primal_weight.
        use_network=False: the training solutions' mean in the reduced basis for
        every instance (the ablation that shows what the network adds)."""
        if use_network:
            C = self.net.forward(self.features(lps)).astype(np.float64) * self.c_std + self.c_mean
        else:
            C = np.tile(self.c_mean, (len(lps), 1))
        Z = self.z_mean + C @ self.z_basis.T
        n = self.base.cols
        out = []
        for lp, z in zip(lps, Z):
            x = z[:n] * self._upper(lp)
            lo = np.where(lp.var_lower["kind"] == mp.FINITE, lp.var_lower["value"], -np.inf)
            hi = np.where(lp.var_upper["kind"] == mp.FINITE, lp.var_upper["value"], np.inf)
            out.append((np.clip(x, lo, hi), z[n:] * self.sy))
        return out

    def solve_batch(self, lps, omega=None, use_network: bool = True, **kw):
        """Warm-started batched solve (mcpdlp.solve_batch keywords) with the
        cold start as the fallback.  Returns (results, fallbacks): an instance
        re-solved cold carries 'warm_fallback': True and the warm attempt's
        iterations added to its own."""
        pred = self.predict(lps, use_network)
        w = self.omega if omega is None else omega
        kw_w = dict(kw)
        if self.cap > 0:
            kw_w["max_iterations"] = min(int(kw.get("max_iterations", self.cap)), self.cap)
        res = mp.solve_batch(lps, warm_primal=np.concatenate([p[0] for p in pred]),
                             warm_dual=np.concatenate([p[1] for p in pred]), initial_primal_weight=w, **kw_w)
        bad = [i for i, r in enumerate(res) if r["status"] != "optimal"]
        if bad:
            cold = mp.solve_batch([lps[i] for i in bad], **kw)
            for i, r in zip(bad, cold):
                r["iterations"] += res[i]["iterations"]
                r["warm_fallback"] = True
                res[i] = r
        return res, len(bad)

    def tune_omega(self, lps, **kw) -> dict:
        """Choose the start primal weight on a validation batch: the training
        solves' median final weight, or the cold-start default (0), whichever
        needs fewer iterations in total (fallbacks included)."""
        tried = {}
        for w in (self.omega, 0.0):
            res, nb = self.solve_batch(lps, omega=w, **kw)
            tried[w] = (sum(r["iterations"] for r in res), nb)
        best = min(tried, key=lambda w: tried[w][0])
        out = {"median_final_weight": self.omega, "iterations_median_weight": tried[self.omega][0],
               "iterations_default_weight": tried[0.0][0], "chosen": "median" if best == self.omega else "default"}
        self.omega = best
        return out
