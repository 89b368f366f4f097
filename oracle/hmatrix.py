"""Restatement of the reference's H-matrix operator (src/HMatrix.f90) in
numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

partition()  -- SPLIT_CLUSTER_ITERATIVE (HMatrix.f90:232-371) and
                BUILD_BLOCK_TREE_ITERATIVE (:379-483), 0-based.
aca()        -- ADD_LOWRANK_BLOCK (:556-750).  The stopping norm uses the
                exact cross terms 2 Re(<u_l, u_k> <v_k, v_l>) of ||U VH||_F^2;
                the reference's update (:703-709) multiplies real parts entry
                by entry over J = 1..NR and reads VH past its NC columns when
                NR > NC, so that step is restated as the criterion it states
                (|u_k| |v_k| < tol ||U VH||_F), not as written.
build()      -- HMATRIX_BUILD (:135-199): ACA blocks kept low-rank only when
                rank (nr + nc) < nr nc (:731-737).
matvec()     -- HMATVEC (:764-820).
"""
from __future__ import annotations

import numpy as np


def partition(cen, leaf=32, eta=2.0):
    """cen: n x 3.  Returns (perm, blocks): perm[k] = panel at cluster
    position k; blocks = list of (rs, nr, cs, nc, aca) in creation order."""
    cen = np.asarray(cen, dtype=np.float64)
    n = cen.shape[0]
    perm = list(range(n))
    w = [list(map(float, cen[i])) for i in range(n)]
    tree = []  # [s, e, left, right, c(3), diam]
    stack = [(0, n - 1, -1, 0)]
    while stack:
        s, e, parent, side = stack.pop()
        ns = e - s + 1
        if ns <= 0:
            continue
        node = len(tree)
        tree.append([s, e, -1, -1, [0.0, 0.0, 0.0], 0.0])
        if parent >= 0:
            tree[parent][2 if side == 1 else 3] = node
        bmin = list(w[s])
        bmax = list(w[s])
        acc = [0.0, 0.0, 0.0]
        for j in range(s, e + 1):
            for d in range(3):
                v = w[j][d]
                if v < bmin[d]:
                    bmin[d] = v
                if v > bmax[d]:
                    bmax[d] = v
                acc[d] += v
        ln = [bmax[d] - bmin[d] for d in range(3)]
        tree[node][4] = [acc[d] / float(ns) for d in range(3)]
        tree[node][5] = float(np.sqrt(ln[0] * ln[0] + ln[1] * ln[1] + ln[2] * ln[2]))
        if ns <= leaf:
            continue
        sd, ml = 0, ln[0]
        if ln[1] > ml:
            sd, ml = 1, ln[1]
        if ln[2] > ml:
            sd = 2
        split = 0.5 * (bmin[sd] + bmax[sd])
        i, j = s, e
        while i <= j:
            if w[i][sd] <= split:
                i += 1
            else:
                w[i], w[j] = w[j], w[i]
                perm[i], perm[j] = perm[j], perm[i]
                j -= 1
        le = i - 1
        if le < s:
            le = s
        if le >= e:
            le = e - 1
        stack.append((le + 1, e, node, 2))
        stack.append((s, le, node, 1))
    blocks = []
    bs = [(0, 0)]
    while bs:
        r, c = bs.pop()
        R, Cn = tree[r], tree[c]
        nr, nc = R[1] - R[0] + 1, Cn[1] - Cn[0] + 1
        if nr <= 0 or nc <= 0:
            continue
        lr, lc = R[2] < 0, Cn[2] < 0
        dx, dy, dz = (R[4][0] - Cn[4][0], R[4][1] - Cn[4][1], R[4][2] - Cn[4][2])
        dist = float(np.sqrt(dx * dx + dy * dy + dz * dz))
        if dist > eta * max(R[5], Cn[5]) and nr > 1 and nc > 1:
            blocks.append((R[0], nr, Cn[0], nc, 1 if (nr > 4 and nc > 4) else 0))
            continue
        if lr and lc:
            blocks.append((R[0], nr, Cn[0], nc, 0))
            continue
        if lr:
            bs += [(r, Cn[3]), (r, Cn[2])]
        elif lc:
            bs += [(R[3], c), (R[2], c)]
        else:
            bs += [(R[3], Cn[3]), (R[3], Cn[2]), (R[2], Cn[3]), (R[2], Cn[2])]
    return np.array(perm, dtype=np.int64), blocks


def aca(Ab, rank_max=50, tol=1e-4):
    """Partially pivoted ACA of the block Ab (nr x nc, complex).  Returns
    U (nr x k), VH (k x nc)."""
    nr, nc = Ab.shape
    kmax = min(rank_max, nr, nc)
    U = np.zeros((nr, kmax), dtype=np.complex128)
    VH = np.zeros((kmax, nc), dtype=np.complex128)
    used = np.zeros(nr, dtype=bool)
    frob = 0.0
    rank = 0
    pr = 0
    for k in range(kmax):
        if k > 0:
            cand = np.where(used, -1.0, np.abs(U[:, k - 1]))
            if cand.max() < 0.0:
                break
            pr = int(np.argmax(cand))
        used[pr] = True
        row = Ab[pr, :] - U[pr, :k] @ VH[:k, :]
        pc = int(np.argmax(np.abs(row)))
        piv = row[pc]
        if abs(piv) < 1e-30:
            break
        col = (Ab[:, pc] - U[:, :k] @ VH[:k, pc]) / piv
        U[:, k] = col
        VH[k, :] = row
        rank = k + 1
        uvn = np.linalg.norm(col) * np.linalg.norm(row)
        frob += uvn * uvn
        if k > 0:
            du = U[:, :k].conj().T @ col            # <u_l, u_k>
            dv = VH[:k, :].conj() @ row             # <v_k, v_l> = sum_j v_k(j) conj(v_l(j))
            frob += 2.0 * float(np.sum((du * dv).real))
        if uvn < tol * np.sqrt(max(frob, 1e-30)):
            break
    return U[:, :rank], VH[:rank, :]


def build(A, cen, leaf=32, eta=2.0, tol=1e-4, rank_max=50):
    """H-matrix of A (n x n): (perm, blocks), each block (rs, nr, cs, nc,
    kind, data): kind 'lr' with (U, VH) or 'dense' with D."""
    perm, parts = partition(cen, leaf, eta)
    out = []
    for rs, nr, cs, nc, is_aca in parts:
        rows, cols = perm[rs:rs + nr], perm[cs:cs + nc]
        Ab = A[np.ix_(rows, cols)]
        if is_aca:
            U, VH = aca(Ab, rank_max, tol)
            k = U.shape[1]
            if k > 0 and k * (nr + nc) < nr * nc:
                out.append((rs, nr, cs, nc, "lr", (U, VH)))
                continue
        out.append((rs, nr, cs, nc, "dense", Ab))
    return perm, out


def matvec(H, x):
    perm, blocks = H
    x = np.asarray(x)
    xp = x[perm]
    yp = np.zeros_like(xp)
    for rs, nr, cs, nc, kind, data in blocks:
        if kind == "lr":
            U, VH = data
            yp[rs:rs + nr] += U @ (VH @ xp[cs:cs + nc])
        else:
            yp[rs:rs + nr] += data @ xp[cs:cs + nc]
    y = np.zeros_like(yp)
    y[perm] = yp
    return y
