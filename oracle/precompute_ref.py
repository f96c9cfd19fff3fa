"""ORACLE -- test infrastructure only (used by tests/ as the checker).

Pure numpy / Python restatement of the host precompute for small meshes,
written independently of the library's C++:

  rcm_order          reverse Cuthill-McKee with a George-Liu pseudo-peripheral
                     start, ties broken by (degree, index) -- the renumbering
                     the reference gets from `renumberMesh` (README.md:76)
  compact_stencils   face neighbours in ascending face order, then boundary
                     pseudo-cells (basicCUDAsolver.cpp:275-299)
  vertex_stencils    edgeNeighbour stencils: compact, then cells sharing a
                     vertex, polygon order then ascending id (basicCUDAsolver.cpp:356-375)
  face_stencils      faceNeighbour stencils (basicCUDAsolver.cpp:333-355)
  rbf_coefficients   reference distance, moment matrix P, pseudo-inverse by
                     numpy.linalg.pinv (LAPACK SVD), face coefficients
                     a = G^T b - mean + 1/n (basicCUDAsolver.cpp:461-619)
"""
from collections import deque

import numpy as np


def rcm_order(n, adj):
    """adj: list of sorted neighbour lists.  Returns new -> old order."""
    deg = [len(a) for a in adj]
    done = [False] * n
    order = []

    def levels(s):
        lev = {s: 0}
        q = deque([s])
        seen = [s]
        while q:
            v = q.popleft()
            for w in adj[v]:
                if done[w] or w in lev:
                    continue
                lev[w] = lev[v] + 1
                q.append(w)
                seen.append(w)
        ecc = lev[seen[-1]]
        last = [v for v in reversed(seen) if lev[v] == ecc]
        # reversed(seen) stops at the first node of another level in the C++;
        # the last level is contiguous at the end of the BFS queue
        return seen, ecc, last

    nxt = 0
    while len(order) < n:
        while done[nxt]:
            nxt += 1
        comp, _, _ = levels(nxt)
        s = min(comp, key=lambda v: (deg[v], v))
        _, ecc, last = levels(s)
        for _ in range(32):
            best = min(last, key=lambda v: (deg[v], v))
            _, ecc2, last2 = levels(best)
            if ecc2 <= ecc:
                break
            s, ecc, last = best, ecc2, last2
        base = len(order)
        order.append(s)
        done[s] = True
        h = base
        while h < len(order):
            v = order[h]
            h += 1
            nb = sorted({w for w in adj[v] if not done[w]}, key=lambda w: (deg[w], w))
            for w in nb:
                done[w] = True
                order.append(w)
    return order[::-1]


def compact_stencils(nc, nf, own, nei, bown):
    cc = [[] for _ in range(nc)]
    cf = [[] for _ in range(nc)]
    for f in range(nf):
        cc[own[f]].append(nei[f])
        cf[own[f]].append(f + 1)
        cc[nei[f]].append(own[f])
        cf[nei[f]].append(-(f + 1))
    for b, c in enumerate(bown):
        cc[c].append(nc + b)
        cf[c].append(nf + b + 1)
    return cc, cf


def vertex_stencils(nc, cc, cell_pts):
    """cell_pts: list of canonical point-id lists per cell (polygon order)."""
    pcells = {}
    for c, pts in enumerate(cell_pts):
        for p in pts:
            pcells.setdefault(p, []).append(c)
    out = []
    for c in range(nc):
        st = list(cc[c])
        seen = set(st) | {c}
        for p in cell_pts[c]:
            for x in sorted(pcells[p]):
                if x not in seen:
                    seen.add(x)
                    st.append(x)
        out.append(st)
    return out


def face_stencils(nc, cc):
    out = []
    for c in range(nc):
        st = list(cc[c])
        seen = set(st) | {c}
        for nb in cc[c]:
            if nb >= nc:
                continue
            for x in cc[nb]:
                if x not in seen:
                    seen.add(x)
                    st.append(x)
        out.append(st)
    return out


def rbf_coefficients(cell, stencil, centres, gauss, face_centre, shape_factor):
    """Coefficients of one face side.

    stencil: member ids (without the cell); centres[id] -> (x, y);
    gauss[id] -> (points (k,2), weights (k,)).  Returns (a, trouble, G)."""
    members = [cell] + list(stencil)
    X = np.array([centres[m] for m in members], dtype=np.float64)
    ref = np.mean(np.hypot(X[1:, 0] - X[0, 0], X[1:, 1] - X[0, 1]))
    n = len(members)
    P = np.zeros((n, n))
    for i, m in enumerate(members):
        g, w = gauss[m]
        d = (g[:, None, :] - X[None, :, :]) / ref
        P[i] = (w[:, None] * np.sqrt(shape_factor + (d ** 2).sum(-1))).sum(0)
    G = np.linalg.pinv(P, rcond=0.0)
    trouble = bool(np.any(np.abs(G) > 1e10))
    if trouble:
        a = np.zeros(n)
        a[0] = 1.0
        return a, True, G
    b = np.sqrt(shape_factor + (((np.asarray(face_centre) - X) / ref) ** 2).sum(-1))
    v = G.T @ b
    return v - v.sum() / n + 1.0 / n, False, G
