"""Shared panels for a batch of distributions over one particle (the reference's batch groups,
scatrs/src/ensemble/engine.rs:257-296 and 725-924): at each wavelength every distribution with
a range is integrated on one set of panels, so that each radius node is solved once for all of
them.

* Pass 1 (``pass1_groups``): a group's panels start from the edges of its members' ranges,
  each interval cut into panels no wider than a quarter of the narrowest member range that
  covers it. Each panel's 15 nodes are solved once (the single-particle batch kernels); every
  member whose range overlaps a panel gets the panel's sums with its own weights. A member's
  targets are the reference's (those of ``adaptive.pass1``), its errors the sums of
  |Kronrod - Gauss| over its panels. A panel is split when its error exceeds a failing
  member's target over that member's panel count (worst first, within ``max_panels`` per
  group, none narrower than 1e-10 in ln r); a member whose outer panel holds more than a tenth
  of a target extends its range by that panel's width, onto the group's next panel or onto a
  new one. The group stops when every member meets its targets and wants no extension, or
  when nothing can be split or added (then the members that miss are unconverged).
* Pass 2 (``pass2_groups``): every node of a group's panels is solved once with its own exact
  angle rule (N_max + (L+1)/2 cosines on the adaptive engine's ladder) into its
  unnormalized scalars, Greek coefficients and phase matrix; each member's sums are its
  weights times those values, in rows of 16 of its panels for the error bounds, by both rules.

The members' results differ from their own unshared runs within rtol (different panels), as
the reference's batch results differ from its single runs.
"""

from __future__ import annotations

import numpy as np

from . import adaptive as A
from .size import MONODISPERSE, POWER_LAW, SEGMENT

__all__ = ["pass1_groups", "pass2_groups"]


def _pair_sums(tasks, pt, pa, pb, V, fold):
    """Kronrod, Gauss and |integrand| Kronrod sums (n, 5) of tasks pt on panels [pa, pb] whose
    15 node values are V (n, 15, 5)."""
    _, w = A._nodes(tasks, pt, pa, pb, np.arange(15))
    K = np.einsum("nk,nkq->nq", w, V)
    G = np.einsum("nk,nkq->nq", w * fold, V)
    S = np.einsum("nk,nkq->nq", w, np.abs(V))
    return K, G, S


def pass1_groups(tasks, groups, rtol: float, max_panels: int, refine_backscatter: bool,
                 allow_unvalidated: bool) -> int:
    """Places the panels of every group (lists of task indices sharing one particle at one
    wavelength); sets each task's panels (a, b), sums (K, G), range and converged flag as
    adaptive.pass1 does, and ``shared_owner``, the task whose particle solves the group's
    nodes. Returns the Mie terms evaluated."""
    A._prepare(tasks, rtol)
    T = len(tasks)
    for q in tasks:
        if q.kind == MONODISPERSE:
            q.done = q.converged = True
        elif q.theta or q.dirs:
            raise ValueError("shared panels refine no derivatives")
    glist = []
    for g in groups:
        live = np.array([t for t in g if tasks[t].kind != MONODISPERSE], np.int64)
        if live.size:
            glist.append(live)
    if not glist:
        return 0
    Gn = len(glist)
    gid = np.full(T, -1)
    for g, mem in enumerate(glist):
        gid[mem] = g
    rep = np.array([mem[0] for mem in glist])
    kinds = np.array([q.kind for q in tasks])
    fixed = np.isin(kinds, (POWER_LAW, SEGMENT))
    t_lo = np.array([q.t_lo for q in tasks])
    t_hi = np.array([q.t_hi for q in tasks])
    pa, pb, pg = [], [], []
    for g, mem in enumerate(glist):
        ends = np.unique(np.concatenate([t_lo[mem], t_hi[mem]]))
        for a0, b0 in zip(ends[:-1], ends[1:]):
            cover = mem[(t_lo[mem] <= a0) & (t_hi[mem] >= b0)]
            if not cover.size:
                continue  # a gap between members' ranges
            w0 = np.min(t_hi[cover] - t_lo[cover]) / A._INITIAL_PANELS
            k = max(1, int(np.ceil((b0 - a0) / w0 - 1e-9)))
            e = np.linspace(a0, b0, k + 1)
            pa.append(e[:-1])
            pb.append(e[1:])
            pg.append(np.full(k, g))
    a, b, grp = np.concatenate(pa), np.concatenate(pb), np.concatenate(pg)
    pid = np.arange(a.size)
    next_id = a.size
    V = np.full((a.size, 15, A._Q), np.nan)
    ck = np.zeros(0, np.int64)  # cached pair sums, by panel id x T + task
    cK = cG = cS = np.zeros((0, A._Q))
    fold = np.where(A._WGF > 0, A._WGF / A._WK, 0.0)[None, :]
    gdone = np.zeros(Gn, bool)
    conv = np.zeros(T, bool)
    final = {}  # task -> (a, b, K, G) of its panels when its group stopped
    terms = 0
    for rnd in range(A._MAX_ROUNDS):
        order = np.lexsort((a, grp))
        a, b, grp, pid, V = a[order], b[order], grp[order], pid[order], V[order]
        pending = np.flatnonzero(np.isnan(V[:, 0, 0]) & ~gdone[grp])
        if pending.size:
            r, _ = A._nodes(tasks, rep[grp[pending]], a[pending], b[pending], np.arange(15))
            q, _, nt = A._node_values(tasks, np.repeat(rep[grp[pending]], 15), r.reshape(-1),
                                      allow_unvalidated)
            terms += nt
            V[pending] = q.reshape(pending.size, 15, A._Q)
        # pairs: each open group's panels with the members whose range overlaps them
        P = a.size
        pp, pt = [], []
        for g, mem in enumerate(glist):
            if gdone[g]:
                continue
            sel = np.flatnonzero(grp == g)
            hit = (a[sel, None] < t_hi[None, mem]) & (b[sel, None] > t_lo[None, mem])
            i, j = np.nonzero(hit)
            pp.append(sel[i])
            pt.append(mem[j])
        pp, pt = np.concatenate(pp), np.concatenate(pt)
        key = pid[pp] * T + pt
        o = np.argsort(key)
        pp, pt, key = pp[o], pt[o], key[o]
        K = np.empty((key.size, A._Q))
        Gs = np.empty((key.size, A._Q))
        S = np.empty((key.size, A._Q))
        have = np.zeros(key.size, bool)
        if ck.size:
            pos = np.minimum(np.searchsorted(ck, key), ck.size - 1)
            have = ck[pos] == key
            K[have], Gs[have], S[have] = cK[pos[have]], cG[pos[have]], cS[pos[have]]
        new = np.flatnonzero(~have)
        if new.size:
            K[new], Gs[new], S[new] = _pair_sums(tasks, pt[new], a[pp[new]], b[pp[new]],
                                                 V[pp[new]], fold)
        ck, cK, cG, cS = key, K, Gs, S
        # per member: totals, errors and targets as adaptive.pass1's (engine.rs:620-716)
        E = np.abs(K - Gs)
        seg = lambda v: np.stack([np.bincount(pt, v[:, j], minlength=T)  # noqa: E731
                                  for j in range(A._Q)], 1)
        tot, esum, asum = seg(K), seg(E), seg(S)
        zero = np.zeros(T, np.int64)
        tgt, end_tgt = A._targets(tot, asum, rtol, 0, 0, zero, zero, refine_backscatter)
        miss = esum > tgt
        fail = miss.any(1)
        n_task = np.bincount(pt, minlength=T)
        first_a = np.full(T, np.inf)
        last_b = np.full(T, -np.inf)
        np.minimum.at(first_a, pt, a[pp])
        np.maximum.at(last_b, pt, b[pp])
        big = (np.abs(K) > 0.1 * end_tgt[pt]).any(1) & ~fixed[pt]
        want_lo = np.zeros(T, bool)
        want_hi = np.zeros(T, bool)
        want_lo[pt[big & (a[pp] == first_a[pt])]] = True
        want_hi[pt[big & (b[pp] == last_b[pt])]] = True
        ok = ~fail & ~want_lo & ~want_hi
        by_task = np.lexsort((a[pp], pt))  # each member's pairs in ln r order
        tb = np.searchsorted(pt[by_task], np.arange(T + 1))

        def finish(g, converged):
            gdone[g] = True
            for t in glist[g]:
                conv[t] = converged[t]
                sel = by_task[tb[t]:tb[t + 1]]
                final[t] = (a[pp[sel]], b[pp[sel]], K[sel], Gs[sel])

        for g, mem in enumerate(glist):
            if not gdone[g] and ok[mem].all():
                finish(g, ok)
        # Within max_panels per group: the members' extensions first, onto the group's next
        # panel where there is one, else onto a new panel of their outer panel's width (one
        # per group and side); then the splits, worst first (engine.rs:881-909).
        room = np.full(Gn, max_panels) - np.bincount(grp, minlength=Gn)
        add_a, add_b, add_g = [], [], []
        progress = np.zeros(Gn, bool)
        made = {}
        for side, want in ((0, want_lo), (1, want_hi)):
            for t in np.flatnonzero(want & (gid >= 0)):
                g = gid[t]
                if gdone[g]:
                    continue
                mine = pp[by_task[tb[t]:tb[t + 1]]]
                j = mine[0] if side == 0 else mine[-1]
                e, w = (a[j], b[j] - a[j]) if side == 0 else (b[j], b[j] - a[j])
                nxt = np.flatnonzero((grp == g) & ((b == e) if side == 0 else (a == e)))
                if nxt.size:
                    new_end = a[nxt[0]] if side == 0 else b[nxt[0]]
                elif (g, side, e) in made:
                    new_end = made[(g, side, e)]
                elif room[g] >= 1:
                    new_end = e - w if side == 0 else e + w
                    made[(g, side, e)] = new_end
                    add_a.append(min(e, new_end))
                    add_b.append(max(e, new_end))
                    add_g.append(g)
                    room[g] -= 1
                else:
                    continue
                if side == 0:
                    t_lo[t] = new_end
                else:
                    t_hi[t] = new_end
                progress[g] = True
        tg = np.where(tgt > 0, tgt, np.finfo(float).tiny)
        ratio = np.where(miss[pt], E / tg[pt], 0.0).max(1) * n_task[pt]
        bad = np.zeros(P)
        np.maximum.at(bad, pp, ratio)
        cand = np.flatnonzero((bad > 1.0) & ~gdone[grp] & (b - a > A.MIN_WIDTH))
        o = cand[np.lexsort((-bad[cand], grp[cand]))]
        og = grp[o]
        rank = np.arange(o.size) - np.searchsorted(og, og)
        cut = np.zeros(P, bool)
        cut[o[rank < room[og]]] = True
        progress[grp[cut]] = True
        for g in range(Gn):
            if not gdone[g] and (not progress[g] or rnd == A._MAX_ROUNDS - 1):
                finish(g, ok)  # out of panels, nothing left to split, or out of rounds
        if gdone.all():
            break
        cut &= ~gdone[grp]
        keep_new = [k for k, g in enumerate(add_g) if not gdone[g]]
        add_a = [add_a[k] for k in keep_new]
        add_b = [add_b[k] for k in keep_new]
        add_g = [add_g[k] for k in keep_new]
        mid = 0.5 * (a + b)
        keep = ~cut
        nc = int(cut.sum())
        ids = np.arange(next_id, next_id + 2 * nc + len(add_g))
        next_id += ids.size
        a = np.concatenate([a[keep], a[cut], mid[cut], np.array(add_a, float)])
        b = np.concatenate([b[keep], mid[cut], b[cut], np.array(add_b, float)])
        grp = np.concatenate([grp[keep], grp[cut], grp[cut], np.array(add_g, np.int64)])
        pid = np.concatenate([pid[keep], ids])
        V = np.concatenate([V[keep], np.full((ids.size, 15, A._Q), np.nan)])
    for g, mem in enumerate(glist):
        for t in mem:
            q = tasks[t]
            q.a, q.b, q.K, q.G = (v.copy() for v in final[t])
            q.t_lo, q.t_hi = float(t_lo[t]), float(t_hi[t])
            q.converged = bool(conv[t])
            q.done = True
            q.shared_owner = int(rep[g])
    return terms


def pass2_groups(tasks, *, order, n_angles, mu, want_phase, allow_unvalidated):
    """Pass 2 of tasks placed by pass1_groups (and monodisperse tasks): each distinct node
    (group panel, Kronrod abscissa) solved once with its own angle rule, the tasks' sums by both
    rules. Returns [(avg (T, 1, 7), greek (T, 1, 6, L+1), phase (T, 1, 6, n), terms, rule,
    rows)] for the Kronrod and the Gauss rule, as adaptive.pass2_torch returns one rule (R = 1)
    (tensors on the current device). A group's row sums are one dense product of its rows'
    weights (rows x the group's nodes, each weight at its own place) with the nodes' values."""
    import torch

    from .ensemble import ensemble_nodes_device, number_averages
    from .mie import default_n_max

    dev = torch.device("cuda", torch.cuda.current_device())
    T = len(tasks)
    W = 0 if order is None else order + 1
    single = want_phase and mu is None
    f64 = torch.float64
    tt = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)  # noqa: E731
    mono = np.array([q.kind == MONODISPERSE for q in tasks])
    owner = np.array([getattr(q, "shared_owner", t) for t, q in enumerate(tasks)], np.int64)
    # every (task, panel) of the non-monodisperse tasks, in task order
    cnt = np.array([0 if m else q.a.size for q, m in zip(tasks, mono)], np.int64)
    p_task = np.repeat(np.arange(T), cnt)
    p_a = np.concatenate([q.a for q, m in zip(tasks, mono) if not m] or [np.zeros(0)])
    p_b = np.concatenate([q.b for q, m in zip(tasks, mono) if not m] or [np.zeros(0)])
    p_own = owner[p_task]
    # distinct blocks of 15 nodes: (owner, a, b); a monodisperse law: one node of its own
    keys = np.stack([p_own.astype(np.float64), p_a, p_b], 1)
    uk, blk = np.unique(keys, axis=0, return_inverse=True) if keys.size else \
        (np.zeros((0, 3)), np.zeros(0, np.int64))
    blk = blk.reshape(-1)
    nb = uk.shape[0]
    half, mid = 0.5 * (uk[:, 2] - uk[:, 1]), 0.5 * (uk[:, 1] + uk[:, 2])
    r_blocks = np.exp(mid[:, None] + half[:, None] * A._X[None, :])  # (nb, 15)
    b_task = np.zeros(nb, np.int64)
    b_task[blk] = p_task  # a task of each block (its group's particle)
    mt = np.flatnonzero(mono)
    r_all = np.concatenate([r_blocks.reshape(-1), np.array([tasks[t].p[0] for t in mt])])
    t_all = np.concatenate([np.repeat(b_task, 15), mt])
    n = r_all.size
    wl_t = np.array([q.wl for q in tasks])
    if order is None:
        rule = np.full(n, 2 if n_angles is None else int(n_angles))
    elif n_angles is not None:
        rule = np.full(n, int(n_angles))
    else:
        need = default_n_max(2.0 * np.pi * r_all / wl_t[t_all]) + (order + 2) // 2
        rule = A._rule_ladder(need).astype(np.int64)
        if single:
            rule[:] = rule.max()
    n_phase = (int(rule[0]) if mu is None else int(np.size(mu))) if want_phase else 0
    Qn = 7 + 6 * W + 6 * n_phase
    F = torch.zeros((n, Qn), dtype=f64, device=dev)
    terms = 0
    layers_t = np.array([q.layers for q in tasks])
    for L in np.unique(layers_t[t_all]):
        for na in np.unique(rule[layers_t[t_all] == L]):
            sel = np.flatnonzero((layers_t[t_all] == L) & (rule == na))
            ts = t_all[sel]
            nodes = np.stack([r_all[sel], np.ones(sel.size)], -1)[:, None, :]
            m = np.stack([tasks[t].m for t in ts])
            fr = np.stack([tasks[t].fr for t in ts]) if L > 1 else None
            sc, g, ph_, nt = ensemble_nodes_device(
                tt(nodes), tt(wl_t[ts]), tt(m[:, 0] if L == 1 else m),
                None if fr is None else tt(fr), n_angles=int(na), order=order,
                normalized=False, allow_unvalidated=allow_unvalidated, want_phase=want_phase,
                mu=mu)
            terms += nt
            cols = [number_averages(sc)]
            if W:
                cols.append(g.reshape(sel.size, -1))
            if n_phase:
                cols.append(ph_.reshape(sel.size, -1))
            F[tt(sel)] = torch.cat(cols, 1)
    # rows: 16 panels of a task (a monodisperse law: its node, weight 1), in task order
    first = np.concatenate([[0], np.cumsum(cnt)])
    p_row_in_task = np.arange(p_task.size) - first[p_task]
    rows_per = np.where(mono, 1, (cnt + A._ROW_PANELS - 1) // A._ROW_PANELS)
    row0 = np.concatenate([[0], np.cumsum(rows_per)])
    p_row = row0[p_task] + p_row_in_task // A._ROW_PANELS
    nr = int(row0[-1])
    row_task = np.repeat(np.arange(T), rows_per)
    _, wk = A._nodes(tasks, p_task, p_a, p_b, np.arange(15)) if p_task.size else \
        (None, np.zeros((0, 15)))
    fold = np.where(A._WGF > 0, A._WGF / A._WK, 0.0)
    # each (row, node) entry: its row, its node's index, its Kronrod weight
    e_row = np.concatenate([np.repeat(p_row, 15), row0[mt]])
    e_node = np.concatenate([(blk[:, None] * 15 + np.arange(15)[None, :]).reshape(-1),
                             nb * 15 + np.arange(mt.size)])
    e_wk = np.concatenate([wk.reshape(-1), np.ones(mt.size)])
    e_wg = np.concatenate([(wk * fold[None, :]).reshape(-1), np.ones(mt.size)])
    # the product per group of rows sharing nodes (a block's owner; monodisperse rows alone)
    e_grp = np.concatenate([np.repeat(p_own, 15), -1 - mt])
    out_rows = [torch.zeros((nr, Qn), dtype=f64, device=dev) for _ in range(2)]
    zrow = [np.zeros(nr), np.zeros(nr)]
    for k, ew in enumerate((e_wk, e_wg)):
        np.add.at(zrow[k], e_row, ew)
    order_e = np.argsort(e_grp, kind="stable")
    bounds = np.flatnonzero(np.diff(e_grp[order_e])) + 1
    for part in np.split(order_e, bounds):
        rows_u, r_loc = np.unique(e_row[part], return_inverse=True)
        nodes_u, n_loc = np.unique(e_node[part], return_inverse=True)
        Fg = F[tt(nodes_u)]
        for k, ew in enumerate((e_wk, e_wg)):
            Wg = torch.zeros((rows_u.size, nodes_u.size), dtype=f64, device=dev)
            Wg[tt(r_loc.reshape(-1)), tt(n_loc.reshape(-1))] = tt(ew[part])
            out_rows[k][tt(rows_u)] = Wg @ Fg
    key = tt(row_task)
    mom = tt(np.array([q.moments for q in tasks]))
    single_node = tt(mono)
    out = []
    for k in range(2):
        rows = out_rows[k]
        tot = _task_sums(rows, row0, T)
        zt = tt(np.add.reduceat(zrow[k], row0[:-1]) if nr else np.zeros(T))
        zm = torch.where(single_node, zt, mom[:, 0])
        lin = tot[:, :7]
        lin = torch.cat([lin[:, :5], torch.where(single_node[:, None], lin[:, 5:],
                                                 mom[:, 1:])], 1) / zm[:, None]
        gr = tot[:, 7:7 + 6 * W].reshape(T, 6, W) / zm[:, None, None] if W else None
        ph = (tot[:, 7 + 6 * W:].reshape(T, 6, n_phase) / zm[:, None, None]) if n_phase \
            else None
        rows_out = dict(key=key, lin=rows[:, :7],
                        gr=rows[:, 7:7 + 6 * W].reshape(nr, 6, W) if W else None,
                        ph=rows[:, 7 + 6 * W:].reshape(nr, 6, n_phase) if n_phase else None,
                        z=zm, n_rows=nr, n_mono=0)
        shape = lambda v: None if v is None else v[:, None]  # noqa: E731
        out.append((shape(lin), shape(gr), shape(ph), terms,
                    int(rule.max()) if (single and n) else (int(rule[0]) if n else 2), rows_out))
    return out


def _task_sums(rows, row0, T: int):
    """Each task's rows (rows row0[t] .. row0[t+1]) added in row order (a fixed order: the sums
    do not depend on scheduling)."""
    import torch

    out = torch.zeros((T,) + tuple(rows.shape[1:]), dtype=rows.dtype, device=rows.device)
    per = np.diff(row0)
    width = int(per.max()) if per.size else 0
    for j in range(width):
        has = np.flatnonzero(per > j)
        idx = torch.from_numpy(row0[has] + j).to(rows.device)
        out[torch.from_numpy(has).to(rows.device)] += rows[idx]
    return out
