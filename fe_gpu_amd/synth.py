"""Seeded synthetic panels for the benchmark workloads (host-side, numpy).

The large panel (150M rows) is generated on the device by the library's own
generator; these numpy generators build the small and medium panels used by the
parity tests and the CPU baseline.  FE ids are int32, 1-based, rows ordered by
the first FE dim (individual / worker) and then by period, as a panel file is.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Proposed coefficients: c1 uses the published nlsw estimates
# (Examples/01TESTS_fe_gpu_small_output.txt:434-438); c2 uses tenure 0.4,
# sick_shock -0.2 and the six extras printed by the reference generator
# (Examples/03TESTS_massive_data_output.txt:2).
BETA_C1 = np.array([-0.0030408, 0.0098533, -0.0004487, 0.1222054, 0.0335883])
BETA_C2 = np.array([0.4, -0.2, -0.2665779136218663, 0.06381245236427056, 0.4631683295023366,
                    0.7596569511243322, 0.9276463627630034, -0.09644476806482337])


@dataclass
class Panel:
    y: np.ndarray
    W: np.ndarray            # n x K (C order rows; use np.asfortranarray for column-major)
    fe: np.ndarray           # n_fe x n int32, 1-based
    names: list
    fe_names: list
    beta_true: np.ndarray
    Z: np.ndarray = field(default=None)

    @property
    def n(self):
        return len(self.y)

    @property
    def G(self):
        return self.fe.max(axis=1).astype(np.int32)


def nlsw_like(seed=101, n_ind=4700, n_years=15, n_occ=13):
    """Config 1: unbalanced individual x year panel with a sticky 13-level occupation."""
    rng = np.random.default_rng(seed)
    n_obs_i = 1 + rng.binomial(n_years - 1, 0.3236, size=n_ind)
    ids, years, occs = [], [], []
    for i in range(n_ind):
        yrs = np.sort(rng.choice(n_years, size=n_obs_i[i], replace=False)) + 1
        occ = np.empty(len(yrs), dtype=np.int32)
        cur = rng.integers(1, n_occ + 1)
        for t in range(len(yrs)):
            if t > 0 and rng.random() < 0.15:
                other = rng.integers(1, n_occ)
                cur = other if other < cur else other + 1
            occ[t] = cur
        ids.append(np.full(len(yrs), i + 1, dtype=np.int32))
        years.append(yrs.astype(np.int32))
        occs.append(occ)
    fid = np.concatenate(ids)
    fyear = np.concatenate(years)
    focc = np.concatenate(occs)
    n = len(fid)
    X = np.column_stack([
        rng.normal(37, 10, n),
        rng.gamma(2.0, 3.0, n),
        rng.poisson(2.0, n).astype(np.float64),
        (rng.random(n) < 0.25).astype(np.float64),
        rng.uniform(0, 25, n),
    ])
    a_id = rng.normal(0, 0.4, n_ind)
    a_occ = rng.normal(0, 0.2, n_occ)
    a_year = rng.normal(0, 0.05, n_years)
    y = X @ BETA_C1 + a_id[fid - 1] + a_occ[focc - 1] + a_year[fyear - 1] + rng.normal(0, 0.3, n)
    fe = np.vstack([fid, focc, fyear]).astype(np.int32)
    return Panel(y, X, fe, ["hours", "tenure", "wks_ue", "union", "ttl_exp"], ["id", "occ", "year"], BETA_C1)


def worker_firm_panel(seed=202, n_workers=20000, n_periods=30, n_firms=4000, switch_p=0.10, beta=BETA_C2):
    """Config-2 shape at any size: workers x periods, sticky firms, 8 regressors."""
    rng = np.random.default_rng(seed)
    W_, T_ = n_workers, n_periods
    firm = np.empty((W_, T_), dtype=np.int32)
    tenure = np.empty((W_, T_), dtype=np.float64)
    firm[:, 0] = rng.integers(1, n_firms + 1, W_)
    tenure[:, 0] = 0.0
    for t in range(1, T_):
        sw = rng.random(W_) < switch_p
        newf = rng.integers(1, n_firms + 1, W_)
        firm[:, t] = np.where(sw, newf, firm[:, t - 1])
        tenure[:, t] = np.where(sw, 0.0, tenure[:, t - 1] + 1.0)
    n = W_ * T_
    worker = np.repeat(np.arange(1, W_ + 1, dtype=np.int32), T_)
    period = np.tile(np.arange(1, T_ + 1, dtype=np.int32), W_)
    a_w = rng.normal(0, 1.0, W_)
    a_f = rng.normal(0, 0.5, n_firms)
    a_t = rng.normal(0, 0.1, T_)
    sick = (rng.random(n) < 0.05).astype(np.float64)
    aw_row = a_w[worker - 1]
    extras = 0.3 * aw_row[:, None] + np.sqrt(0.91) * rng.normal(0, 1, (n, 6))
    X = np.column_stack([tenure.reshape(-1), sick, extras])
    f = firm.reshape(-1)
    y = X @ beta[: X.shape[1]] + aw_row + a_f[f - 1] + a_t[period - 1] + rng.normal(0, 1.0, n)
    fe = np.vstack([worker, f, period]).astype(np.int32)
    names = ["tenure", "sick_shock"] + [f"extra_{j}" for j in range(1, 7)]
    return Panel(y, X, fe, names, ["worker", "firm", "time"], beta)


def serving_panel(seed=505, n_workers=500_000, n_periods=20, n_firms=100_000, n_regions=50, n_cand=20,
                  switch_p=0.10):
    """Config-5 serving panel: workers x periods, sticky firms, 20 periods, a 50-level region
    assigned per firm (nested in firm), y plus n_cand candidate columns.  FE dims: worker, firm,
    time, region.  Candidate j = 0.3 alpha_w + 0.2 alpha_f + sqrt(0.87) e_j; y loads on the first
    six candidates."""
    rng = np.random.default_rng(seed)
    W_, T_ = n_workers, n_periods
    sw = rng.random((W_, T_)) < switch_p
    sw[:, 0] = True
    draws = rng.integers(1, n_firms + 1, (W_, T_), dtype=np.int32)
    idx = np.maximum.accumulate(np.where(sw, np.arange(T_)[None, :], 0), axis=1)
    firm = np.take_along_axis(draws, idx, axis=1).reshape(-1)
    n = W_ * T_
    worker = np.repeat(np.arange(1, W_ + 1, dtype=np.int32), T_)
    period = np.tile(np.arange(1, T_ + 1, dtype=np.int32), W_)
    region_of_firm = rng.integers(1, n_regions + 1, n_firms, dtype=np.int32)
    region_of_firm[rng.integers(0, n_firms)] = n_regions
    region = region_of_firm[firm - 1]
    a_w = rng.normal(0, 1.0, W_)[worker - 1]
    a_f = rng.normal(0, 0.5, n_firms)[firm - 1]
    a_t = rng.normal(0, 0.1, T_)[period - 1]
    a_r = rng.normal(0, 0.2, n_regions)[region - 1]
    X = np.empty((n, n_cand), order="F")
    for j in range(n_cand):
        X[:, j] = 0.3 * a_w + 0.2 * a_f + np.sqrt(0.87) * rng.normal(0, 1, n)
    beta = np.zeros(n_cand)
    beta[:6] = [0.5, -0.25, 0.1, 0.8, -0.6, 0.3]
    y = X[:, :6] @ beta[:6] + a_w + a_f + a_t + a_r + rng.normal(0, 1.0, n)
    fe = np.vstack([worker, firm, period, region]).astype(np.int32)
    return Panel(y, X, fe, [f"cand_{j}" for j in range(1, n_cand + 1)], ["worker", "firm", "time", "region"], beta)


def random_fe_problem(seed, n=200, K=3, L=0, G=(7, 5), noise=0.1):
    """Small random problem with arbitrary (unsorted) FE ids for parity tests."""
    rng = np.random.default_rng(seed)
    fe = np.vstack([rng.integers(1, g + 1, n) for g in G]).astype(np.int32)
    for d, g in enumerate(G):          # make max id == g so G_d = g exactly
        fe[d, rng.integers(0, n)] = g
    X = rng.normal(size=(n, K))
    alpha = [rng.normal(size=g) for g in G]
    beta = rng.normal(size=K)
    y = X @ beta + sum(a[fe[d] - 1] for d, a in enumerate(alpha)) + noise * rng.normal(size=n)
    Z = rng.normal(size=(n, L)) + (X[:, :1] if L and K else 0.0) if L else None
    return Panel(y, X, fe, [f"x{j}" for j in range(K)], [f"fe{d}" for d in range(len(G))], beta, Z)
