"""CPU restatement of the reference's H-matrix (TEST INFRASTRUCTURE ONLY).

HMATRIX_BUILD / HMATVEC of HMatrix.f90 in numpy, for small n: the cluster tree
by midpoint bisection of the longest bounding-box side (SPLIT_CLUSTER_ITERATIVE
:225-363), the block partition with eta-admissibility (BUILD_BLOCK_TREE_ITERATIVE
:372-497), the partially pivoted cross approximation of the admissible blocks
with the reference's pivot rules, norm estimate and dense fallbacks
(ADD_LOWRANK_BLOCK :560-748) and the product (HMATVEC :760-820). Only tests/
use it; the product library never imports it.
"""
import numpy as np

NLEAF = 32      # NLEAF_DEFAULT, HMatrix.f90:57
ETA = 2.0       # ETA_DEFAULT, :58
ACA_TOL = 1e-4  # :59
RANK_MAX = 50   # :60


def cluster_tree(xyz, leaf=NLEAF):
    """(perm, nodes): perm[k] = point at cluster position k (0-based); nodes as
    dicts with s, e (positions [s, e)), left, right (-1 for a leaf), c, diam,
    in the order SPLIT_CLUSTER_ITERATIVE creates them (:247-361)"""
    xyz = np.asarray(xyz, dtype=np.float64)
    n = len(xyz)
    w = xyz.copy()
    perm = np.arange(n)
    nodes = []
    stack = [(0, n - 1, -1, 0)]                      # inclusive range, parent, side
    while stack:
        s, e, parent, side = stack.pop()
        ns = e - s + 1
        if ns <= 0:
            continue
        nid = len(nodes)
        if parent >= 0:
            nodes[parent]["left" if side == 1 else "right"] = nid
        pts = w[s:e + 1]
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        ln = hi - lo
        node = dict(s=s, e=e + 1, left=-1, right=-1, c=None,
                    diam=float(np.sqrt(ln[0] * ln[0] + ln[1] * ln[1] + ln[2] * ln[2])))
        # the centre: the coordinates summed in point order (:289-307)
        acc = np.zeros(3)
        for j in range(s, e + 1):
            acc += w[j]
        node["c"] = acc / ns
        nodes.append(node)
        if ns <= leaf:
            continue
        dim = 0
        if ln[1] > ln[dim]:
            dim = 1
        if ln[2] > ln[dim]:
            dim = 2
        split = 0.5 * (lo[dim] + hi[dim])
        i, j = s, e
        while i <= j:                                 # :322-339
            if w[i, dim] <= split:
                i += 1
            else:
                w[[i, j]] = w[[j, i]]
                perm[[i, j]] = perm[[j, i]]
                j -= 1
        left_end = i - 1
        left_end = max(left_end, s)
        left_end = min(left_end, e - 1)
        stack.append((left_end + 1, e, nid, 2))
        stack.append((s, left_end, nid, 1))
    return perm, nodes


def block_partition(nodes, eta=ETA):
    """blocks (rs, nr, cs, nc, admissible-and-compressible) in the order of
    BUILD_BLOCK_TREE_ITERATIVE (:394-493); admissible blocks with a side of at
    most 4 points are stored dense (ADD_LOWRANK_BLOCK :583-586)"""
    out = []
    stack = [(0, 0)]
    while stack:
        r, c = stack.pop()
        a, b = nodes[r], nodes[c]
        nr, nc = a["e"] - a["s"], b["e"] - b["s"]
        if nr <= 0 or nc <= 0:
            continue
        d = a["c"] - b["c"]
        dist = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
        if dist > eta * max(a["diam"], b["diam"]) and nr > 1 and nc > 1:
            out.append((a["s"], nr, b["s"], nc, 0 if (nr <= 4 or nc <= 4) else 1))
            continue
        leaf_r, leaf_c = a["left"] < 0, b["left"] < 0
        if leaf_r and leaf_c:
            out.append((a["s"], nr, b["s"], nc, 0))
        elif leaf_r:
            stack += [(r, b["right"]), (r, b["left"])]
        elif leaf_c:
            stack += [(a["right"], c), (a["left"], c)]
        else:
            stack += [(a["right"], b["right"]), (a["right"], b["left"]), (a["left"], b["right"]),
                      (a["left"], b["left"])]
    return out


def aca(blk, tol=ACA_TOL, kmax=RANK_MAX):
    """ADD_LOWRANK_BLOCK's cross approximation of the dense block blk (nr x nc):
    returns (U, VH) with blk ~= U VH, or None where the reference stores the
    block dense (the rank does not pay, or rank 0; :728-734)"""
    nr, nc = blk.shape
    km = min(kmax, nr, nc)
    U = np.zeros((nr, km), dtype=complex)
    VH = np.zeros((km, nc), dtype=complex)
    used = np.zeros(nr, dtype=bool)
    frob = 0.0
    rank = 0
    for k in range(km):
        if k == 0:
            piv = 0
        else:
            cand = np.where(~used)[0]
            if len(cand) == 0:
                break
            piv = int(cand[np.argmax(np.abs(U[cand, k - 1]))])   # first maximum
        used[piv] = True
        row = blk[piv, :] - U[piv, :k] @ VH[:k, :]
        pc = int(np.argmax(np.abs(row)))
        pe = row[pc]
        if abs(pe) < 1e-30:
            break
        col = blk[:, pc] - U[:, :k] @ VH[:k, pc]
        U[:, k] = col / pe
        VH[k, :] = row
        rank = k + 1
        uv = np.linalg.norm(U[:, k]) * np.linalg.norm(VH[k, :])
        frob += uv ** 2
        m = min(nr, nc)
        for i in range(k):
            # the reference's cross term: products of real parts, one index J
            # running over the rows of U and the columns of VH (:699-707; for
            # nr > nc it reads past VH's columns, taken here as the first
            # min(nr, nc) terms)
            frob += 2.0 * np.sum((np.conj(U[:m, i]) * U[:m, k]).real * (np.conj(VH[k, :m]) * VH[i, :m]).real)
        if uv < tol * np.sqrt(max(frob, 1e-30)):
            break
    if rank == 0 or rank * (nr + nc) >= nr * nc:
        return None
    return U[:, :rank], VH[:rank, :]


def build(a, xyz, leaf=NLEAF, eta=ETA, tol=ACA_TOL, kmax=RANK_MAX):
    """HMATRIX_BUILD(N, AMAT, XYZ_P): (perm, stored blocks) with each stored
    block (rs, nr, cs, nc, D or (U, VH)) in cluster positions"""
    perm, nodes = cluster_tree(xyz, leaf)
    ap = np.asarray(a)[np.ix_(perm, perm)]
    blocks = []
    for rs, nr, cs, nc, lr in block_partition(nodes, eta):
        sub = ap[rs:rs + nr, cs:cs + nc]
        f = aca(sub, tol, kmax) if lr else None
        blocks.append((rs, nr, cs, nc, sub.copy() if f is None else f))
    return perm, blocks


def matvec(perm, blocks, x):
    """HMATVEC: y = H x for x of n rows (one or more columns)"""
    x = np.asarray(x, dtype=complex)
    xp = x[perm]
    yp = np.zeros_like(xp)
    for rs, nr, cs, nc, f in blocks:
        if isinstance(f, tuple):
            yp[rs:rs + nr] += f[0] @ (f[1] @ xp[cs:cs + nc])
        else:
            yp[rs:rs + nr] += f @ xp[cs:cs + nc]
    y = np.zeros_like(yp)
    y[perm] = yp
    return y
