"""CPU restatement of the reference's smoothed-aggregation AMG set-up
(Make_amg_matrix.f90), serial semantics.  TEST INFRASTRUCTURE ONLY: used by
tests/ to check the product's host set-up and GPU hierarchy; never shipped.

Pure Python over numpy arrays (sizes up to a few thousand unknowns run in a
second).  Indices are 0-based here; the CSR convention is the reference's
(offind[0] = 0, offind[i+1] = end of row i).

Pinned by: the reference build's dumps (tests/golden/ref_amg*.npz, bit for
bit), a hand-enumerated 1-D chain, and the dense identities P = T - w D~^-1 A_F T,
R = P^T, A_c = R A P (tests/test_oracle.py).
"""
from __future__ import annotations

import numpy as np


def grapfilt(offind, offrow, dmat, lumat, theta):
    """make_grapfilt (Make_amg_matrix.f90:67-111): strong entries (CSR
    positions, in row order) and the filtered diagonal d~ = d + sum(weak)."""
    n = len(dmat)
    th2 = float(theta) * float(theta)
    strong = [[] for _ in range(n)]
    dfil = np.array(dmat, dtype=np.float64)
    for i in range(n):
        for k in range(offind[i], offind[i + 1]):
            j = offrow[k]
            dd = dmat[i] * dmat[j]
            lu = lumat[k]
            if lu * lu > abs(dd * th2) and dd * lu > 0.0:          # :98
                strong[i].append(k)
            else:
                dfil[i] = dfil[i] + lu                               # :102
    return strong, dfil


class _Queue:
    """The aggregation's binary max-heap (Make_amg_matrix.f90:683-768):
    positions 1..size, priority 0 marks a hole; `where[node]` is the node's
    position (0 = not queued)."""

    def __init__(self, members, n):
        self.size = len(members)
        self.prio = [0] * (self.size + 2)
        self.node = [0] * (self.size + 2)
        self.where = [0] * (n + 1)       # 1-based node ids
        for q, g in enumerate(members, start=1):
            self.prio[q], self.node[q], self.where[g] = 1, g, q

    def drop(self, g):                   # remove_que :683-728
        q = self.where[g]
        self.prio[q] = self.node[q] = 0
        c = 2 * q
        while c <= self.size:
            if c + 1 <= self.size and self.prio[c + 1] > self.prio[c]:
                c += 1
            if self.prio[c] == 0:
                break
            moved = self.node[c]
            self.prio[q], self.node[q] = self.prio[c], moved
            self.where[moved] = q
            self.prio[c] = self.node[c] = 0
            q, c = c, 2 * c

    def bump(self, g):                   # priority += 1, then enter_que :730-768
        q = self.where[g]
        self.prio[q] += 1
        while q > 1 and self.prio[q // 2] < self.prio[q]:
            p = q // 2
            self.prio[p], self.prio[q] = self.prio[q], self.prio[p]
            a, b = self.node[q], self.node[p]
            self.node[p], self.node[q] = a, b
            self.where[b], self.where[a] = q, p
            q = p


def aggregate(n, strong_cols):
    """make_aggregation (Make_amg_matrix.f90:113-222): returns the aggregate
    number of every node (1-based aggregates, 0 = none) and the count.
    strong_cols[i] = 0-based columns of node i's strong entries."""
    col = [None] + [[c + 1 for c in s] for s in strong_cols]   # 1-based view
    q = _Queue([i for i in range(1, n + 1) if col[i]], n)
    agg = [0] * (n + 1)
    owner = [0] * (n + 1)                # glob_num
    na = 0
    while q.prio[1] != 0:
        i = q.node[1]
        na += 1
        agg[i] = na
        for j in col[i]:
            agg[j] = na
            if q.where[j] > 0:
                q.drop(j)
            q.where[j] = 0
            owner[j] = i
        q.drop(i)
        q.where[i] = 0
        owner[i] = i
        ring2 = []                       # remove_enter_quedeep :633-681
        for j in col[i]:
            for m in col[j]:
                if owner[m] != i:
                    owner[m] = i
                    ring2.append(m)
                    if q.where[m] > 0:
                        q.drop(m)
                        q.where[m] = 0
        for j in ring2:
            for m in col[j]:
                if owner[m] != i:
                    owner[m] = i
                    if q.where[m] > 0:
                        q.bump(m)
    # orphan check (:173-186): the first node out of the queue that has an
    # aggregated strong neighbour joins it, and the check ends there
    for i in range(1, n + 1):
        if q.where[i] != 0:
            continue
        hit = next((j for j in col[i] if agg[j] > 0), None)
        if hit is not None:
            agg[i] = agg[hit]
            break
    return np.array(agg[1:], dtype=np.int64), na


def complete_orphans(agg, strong_cols):
    """The device AMG's completion of the aggregation (no reference
    counterpart; the reference's fix-up stops after one node,
    Make_amg_matrix.f90:173-186): every node outside an aggregate with an
    aggregated strong neighbour joins the first such neighbour's aggregate,
    in node order, until nothing changes.  `agg` is 0-based by node with
    1-based aggregate numbers (0 = none), as `aggregate` returns."""
    agg = np.array(agg, copy=True)
    changed = True
    while changed:
        changed = False
        for i in range(len(agg)):
            if agg[i] == 0:
                hit = next((j for j in strong_cols[i] if agg[j] > 0), None)
                if hit is not None:
                    agg[i] = agg[hit]
                    changed = True
    return agg


def project(offind, dmat, lumat):
    """The device AMG's M-matrix projection (no reference counterpart):
    couplings max(a_ij, 0), diagonal -max(|a_ii|, sum of the clipped
    couplings).  Pinned by: an M-matrix with a dominant diagonal is left
    unchanged, the result always is one (tests/test_amg.py)."""
    n = len(dmat)
    rows = np.repeat(np.arange(n), np.diff(offind))
    lp = np.maximum(np.asarray(lumat, np.float64), 0.0)
    s = np.zeros(n)
    for k in range(len(lp)):          # row order, as the device sums
        s[rows[k]] += lp[k]
    return -np.maximum(np.abs(np.asarray(dmat, np.float64)), s), lp


def hierarchy(offind, offrow, dmat, lumat, max_levels, theta, omega, orphans=True,
              coarse_max=1):
    """The device AMG's hierarchy in FP64: per level the reference's strength
    filter, heap aggregation (completed when `orphans`) and Jacobi-smoothed
    prolongator, and the Galerkin operator P^T A P (the reference's
    make_galerkin, entry order included).  Returns (mats, prolongs) as scipy
    CSR matrices for `vcycle`, and per coarse level the reference-layout
    arrays; the reference's CSR layout in."""
    import scipy.sparse as sp
    theta, omega = np.float32(theta), np.float32(omega)
    n = len(dmat)
    rows = np.repeat(np.arange(n), np.diff(offind))
    A = (sp.csr_matrix((lumat, (rows, offrow)), shape=(n, n)) + sp.diags(dmat)).tocsr()
    mats, Ps, levels = [A], [], []
    oi, oc, d, lu = np.asarray(offind), np.asarray(offrow), np.asarray(dmat), np.asarray(lumat)
    for lev in range(2, max_levels + 1):
        nn = len(d)
        if nn <= coarse_max:
            break
        strong, dfil = grapfilt(oi, oc, d, lu, theta)
        scols = [[int(oc[k]) for k in s] for s in strong]
        agg, na = aggregate(nn, scols)
        if na == 0 or na >= nn:
            break
        if orphans:
            agg = complete_orphans(agg, scols)
        pi, pc, pv = prolongator(oi, oc, lu, strong, dfil, agg, na, lev, omega)
        P = sp.csr_matrix((pv, pc, pi), shape=(nn, na))
        # the coarse operator in the reference's entry order, which the next
        # aggregation's tie-breaking follows
        ri, rcol, rv = transpose(pi, pc, pv, na)
        oi, oc, d, lu = galerkin(oi, oc, d, lu, ri, rcol, rv)
        rows = np.repeat(np.arange(na), np.diff(oi))
        Ac = (sp.csr_matrix((lu, (rows, oc)), shape=(na, na)) + sp.diags(d)).tocsr()
        Ps.append(P)
        mats.append(Ac)
        levels.append(dict(n=na, offind=oi, offrow=oc, dmat=d, lumat=lu, pindex=pi, poffrow=pc,
                           pval=pv))
    return mats, Ps, levels


def prolongator(offind, offrow, lumat, strong, dfil, agg, na, lev, omega):
    """make_prolomat (Make_amg_matrix.f90:224-371): per fine row the
    aggregates of its strong neighbours (first-appearance order), then its
    own; values -w * sum(lu)/d~ and +(1 - w) on the own aggregate.
    Returns CSR (pindex, 0-based aggregate columns, values)."""
    n = len(agg)
    width = min(30 + 80 * (lev - 2), na)
    omega = float(omega)
    cols, vals = [], []
    for i in range(n):
        c, v = [], []
        for k in strong[i]:
            a = agg[offrow[k]]
            if a and a not in c:
                if len(c) >= width:
                    raise ValueError("prolongation width exceeded")
                c.append(a)
                v.append(0.0)
        if agg[i] and agg[i] not in c:
            if len(c) >= width:
                raise ValueError("prolongation width exceeded")
            c.append(agg[i])
            v.append(0.0)
        if c:
            for k in strong[i]:
                a = agg[offrow[k]]
                if a:
                    s = c.index(a)
                    v[s] = v[s] - lumat[k]
        v = [x * omega / dfil[i] for x in v]
        if agg[i]:
            s = c.index(agg[i])
            v[s] = v[s] + 1.0 - omega
        cols.append(c)
        vals.append(v)
    pindex = np.zeros(n + 1, np.int32)
    pindex[1:] = np.cumsum([len(c) for c in cols])
    pcol = np.array([a - 1 for c in cols for a in c], np.int32)
    pval = np.array([x for v in vals for x in v], np.float64)
    return pindex, pcol, pval


def transpose(pindex, pcol, pval, na):
    """trans_prolo2restr (Make_amg_matrix.f90:373-425): R = P^T, rows in
    aggregate order, each row's fine columns ascending."""
    n = len(pindex) - 1
    rows = [[] for _ in range(na)]
    for i in range(n):
        for k in range(pindex[i], pindex[i + 1]):
            rows[pcol[k]].append((i, pval[k]))
    rindex = np.zeros(na + 1, np.int32)
    rindex[1:] = np.cumsum([len(r) for r in rows])
    rcol = np.array([i for r in rows for i, _ in r], np.int32)
    rval = np.array([v for r in rows for _, v in r], np.float64)
    return rindex, rcol, rval


def galerkin(offind, offrow, dmat, lumat, rindex, rcol, rval):
    """make_galerkin (Make_amg_matrix.f90:427-631), by its result: every
    coarse entry (a, b) is sum over fine f in R row b (ascending) of
    R(b, f) * RA(a, f), RA(a, f) = sum over fine g in R row a (ascending) of
    R(a, g) * A(g, f) (diagonal first, then the CSR row of g), each
    partial sum started at zero as in the reference's scratch arrays.  Rows
    of the coarse CSR list their off-diagonal columns ascending; the
    coarse diagonal defaults to 1 where RAP has no diagonal entry."""
    na = len(rindex) - 1
    ra = []                              # per coarse a: {fine f: value}
    for a in range(na):
        acc = {}
        for k in range(rindex[a], rindex[a + 1]):
            g, r = rcol[k], rval[k]
            acc[g] = acc.get(g, 0.0) + r * dmat[g]
            for q in range(offind[g], offind[g + 1]):
                f = offrow[q]
                acc[f] = acc.get(f, 0.0) + r * lumat[q]
        ra.append(acc)
    # coarse rows b whose R row meets fine f
    meets = {}
    for b in range(na):
        for k in range(rindex[b], rindex[b + 1]):
            meets.setdefault(int(rcol[k]), []).append(b)
    dm = np.ones(na)
    rows = [[] for _ in range(na)]
    for a in range(na):
        cand = sorted({b for f in ra[a] for b in meets.get(int(f), ())})
        for b in cand:
            val, hit = 0.0, False
            for k in range(rindex[b], rindex[b + 1]):
                f = rcol[k]
                if f in ra[a]:
                    val = val + rval[k] * ra[a][f]
                    hit = True
            if hit:
                if a == b:
                    dm[a] = val
                else:
                    rows[a].append((b, val))
    offind_c = np.zeros(na + 1, np.int32)
    offind_c[1:] = np.cumsum([len(r) for r in rows])
    offrow_c = np.array([b for r in rows for b, _ in r], np.int32)
    lumat_c = np.array([v for r in rows for _, v in r], np.float64)
    return offind_c, offrow_c, dm, lumat_c


def make_amgmat(offind, offrow, dmat, lumat, amg_nlevel, amg_theta, jac_omega):
    """make_amgmat (Make_amg_matrix.f90:20-65): levels 1..nlevel as dicts
    with the same keys as the reference dump (0-based column ids)."""
    theta, omega = np.float32(amg_theta), np.float32(jac_omega)
    levels = [dict(n=len(dmat), nnz=int(offind[-1]), offind=np.asarray(offind, np.int32),
                   offrow=np.asarray(offrow, np.int32), dmat=np.asarray(dmat, np.float64),
                   lumat=np.asarray(lumat, np.float64))]
    for lev in range(2, amg_nlevel + 1):
        f = levels[-1]
        if f["n"] == 0:
            break
        strong, dfil = grapfilt(f["offind"], f["offrow"], f["dmat"], f["lumat"], theta)
        agg, na = aggregate(f["n"], [[int(f["offrow"][k]) for k in s] for s in strong])
        pindex, pcol, pval = prolongator(f["offind"], f["offrow"], f["lumat"], strong, dfil,
                                         agg, na, lev, omega)
        rindex, rcol, rval = transpose(pindex, pcol, pval, na)
        oc, rc, dc, lc = galerkin(f["offind"], f["offrow"], f["dmat"], f["lumat"], rindex, rcol,
                                  rval)
        levels.append(dict(n=na, nnz=int(oc[-1]), offind=oc, offrow=rc, dmat=dc, lumat=lc,
                           pindex=pindex, poffrow=pcol, pval=pval, rindex=rindex, roffrow=rcol,
                           rval=rval, aggregate=agg))
    return levels


def vcycle(mats, prolongs, r, npre=2, npost=2, ws=0.67, coarse_sweeps=None):
    """One V-cycle z = M^-1 r of the device AMG preconditioner, restated in
    FP64 numpy: damped-Jacobi sweeps with weight `ws` starting from zero,
    residual, restriction by P^T, the next level, prolongation x += P x_c,
    post-sweeps; the coarsest level is solved exactly, or with
    `coarse_sweeps` Jacobi sweeps from zero.  `mats[l]` is level l's matrix
    (scipy sparse, level 0 = the box), `prolongs[l]` the prolongator from
    level l to l+1.  This is the textbook multigrid V-cycle (Briggs, Henson
    and McCormick, A Multigrid Tutorial, 2nd ed., ch. 3 and 8); the
    reference's own V-cycle is a no-op in serial runs
    (Linear_solution.f90:688-718,820) and has no counterpart here.

    Pinned by: a single level with an exact solve returns A^-1 r; two levels
    without smoothing return the coarse-grid correction P (P^T A P)^-1 P^T r;
    the error propagator I - M^-1 A of a Laplacian contracts
    (tests/test_amg.py)."""
    last = len(mats) - 1

    def jac(A, d, b, x, k):
        for _ in range(k):
            x = x + ws * (b - A @ x) / d
        return x

    def cyc(lv, b):
        A = mats[lv]
        d = A.diagonal()
        if lv == last:
            if coarse_sweeps is None:
                return np.linalg.solve(A.toarray(), b)
            return jac(A, d, b, np.zeros_like(b), coarse_sweeps)
        x = jac(A, d, b, np.zeros_like(b), npre)
        xc = cyc(lv + 1, prolongs[lv].T @ (b - A @ x))
        x = x + prolongs[lv] @ xc
        return jac(A, d, b, x, npost)
    return cyc(0, np.asarray(r, dtype=np.float64))
