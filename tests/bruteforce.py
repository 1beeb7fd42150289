"""Independent dense-algebra restatements used to pin the oracle.

These compute the same quantities as the oracle by a different route: explicit
dummy-variable matrices, least squares through numpy/scipy LAPACK, np.unique for
cluster partitions and numpy.linalg.eigh for the PSD clip.  A dropped term, a
wrong sign or a transposed operand in the oracle's iterative/AP/Cholesky code
shows up as a mismatch against them.
"""
import numpy as np


def dummies(fe):
    """Dense dummy matrix [D_1 | D_2 | ...] for dim-major 1-based ids (empty groups give zero columns)."""
    n = fe.shape[1]
    blocks = []
    for ids in fe:
        G = int(ids.max())
        D = np.zeros((n, G))
        D[np.arange(n), ids - 1] = 1.0
        blocks.append(D)
    return np.hstack(blocks) if blocks else np.zeros((n, 0))


def annihilate(x, fe):
    """M_F x = x - D pinv(D) x (exact within-transform)."""
    D = dummies(fe)
    if D.shape[1] == 0:
        return x.copy()
    coef, *_ = np.linalg.lstsq(D, x, rcond=None)
    return x - D @ coef


def cluster_partition(fe, dims):
    """Dense 0-based labels of the intersection of 1-based dims, plus count (np.unique)."""
    key = np.stack([fe[d - 1] for d in dims], axis=1)
    _, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    return inv, int(inv.max()) + 1


def cgm_meat(labels, nclus, r, X):
    S = np.zeros((nclus, X.shape[1]))
    np.add.at(S, labels, r[:, None] * X)
    return S.T @ S


def fit(y, X, fe, cluster=(), Z=None, endog=()):
    """Brute-force FE regression with the product's (GPU-path) conventions."""
    n, K = X.shape
    yt = annihilate(y, fe)
    Xt = annihilate(X, fe)
    G = fe.max(axis=1)
    dof_fes = int(np.sum(np.maximum(G - 1, 0)))
    endog = [e - 1 for e in endog]
    if endog and Z is not None:
        Zt = annihilate(Z, fe)
        exog = [k for k in range(K) if k not in endog]
        Za = np.hstack([Xt[:, exog], Zt])
        Pz_X = Za @ np.linalg.lstsq(Za, Xt, rcond=None)[0]    # X_hat = P_Z X
        beta = np.linalg.solve(Pz_X.T @ Xt, Pz_X.T @ yt)
        A = Pz_X.T @ Pz_X
        Xs = Pz_X
    else:
        beta = np.linalg.lstsq(Xt, yt, rcond=None)[0]
        A = Xt.T @ Xt
        Xs = Xt
    r = yt - Xt @ beta
    rss = float(r @ r)
    Ainv = np.linalg.inv(A)
    if not cluster:
        df = max(1, n - (K + dof_fes + 1))
        V = rss / df * Ainv
        dof_resid = df
    else:
        dims = list(dict.fromkeys(cluster))[:3]
        V = np.zeros((K, K))
        gmin = None
        m = len(dims)
        for mask in range(1, 1 << m):
            sub = [dims[j] for j in range(m) if mask >> j & 1]
            if len(sub) == 1:
                labels = fe[sub[0] - 1] - 1
                nclus = int(G[sub[0] - 1])
                gmin = nclus if gmin is None else min(gmin, nclus)
            else:
                labels, nclus = cluster_partition(fe, sub)
            w = 1.0 if len(sub) % 2 == 1 else -1.0
            V += w * Ainv @ cgm_meat(labels, nclus, r, Xs) @ Ainv
        nested = 1 if sum(max(int(G[d - 1]) - 1, 0) for d in dims) > 0 else 0
        scalar = (n - 1) / max(1, n - (K + nested))
        if gmin > 1:
            scalar *= gmin / (gmin - 1)
        V = V * scalar
        V = np.triu(V) + np.triu(V, 1).T
        lam, Q = np.linalg.eigh(V)
        V = (Q * np.clip(lam, 0, None)) @ Q.T
        dof_resid = max(1, gmin - 1)
    return dict(beta=beta, se=np.sqrt(np.clip(np.diag(V), 0, None)), V=V, rss=rss, yt=yt, Xt=Xt,
                dof_fes=dof_fes, dof_resid=dof_resid)
