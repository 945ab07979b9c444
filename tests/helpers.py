"""Shared checks for the parity tests."""
import numpy as np


def clusters(w, tol):
    """Group ascending eigenvalues whose neighbours are closer than tol."""
    groups, start = [], 0
    for i in range(1, len(w) + 1):
        if i == len(w) or w[i] - w[i - 1] > tol:
            groups.append((start, i))
            start = i
    return groups


def compare_eig(H_upper, w, V, w_ref, V_ref, hnorm=None):
    """V, V_ref: eigenvectors as rows. Returns dict of error measures."""
    n = H_upper.shape[0]
    H = np.triu(H_upper) + np.triu(H_upper, 1).T
    if hnorm is None:
        hnorm = max(np.max(np.abs(w_ref)), 1e-300)
    out = {}
    out["eig"] = float(np.max(np.abs(w - w_ref)) / hnorm) if n else 0.0
    X = V.T  # columns = eigenvectors
    out["res"] = float(np.max(np.abs(H @ X - X * w[None, :])) / hnorm) if n else 0.0
    out["orth"] = float(np.max(np.abs(X.T @ X - np.eye(n)))) if n else 0.0
    # projectors onto clusters of (near) degenerate eigenvalues, scaled by the gap to the rest
    worst = 0.0
    Xr = V_ref.T
    for (a, b) in clusters(w_ref, 1e-6 * hnorm):
        left = w_ref[a] - w_ref[a - 1] if a > 0 else np.inf
        right = w_ref[b] - w_ref[b - 1] if b < n else np.inf
        gap = min(left, right)
        P = X[:, a:b] @ X[:, a:b].T
        Pr = Xr[:, a:b] @ Xr[:, a:b].T
        err = np.max(np.abs(P - Pr))
        worst = max(worst, err * min(gap, hnorm) / hnorm)
    out["proj"] = float(worst)
    return out
