"""A host emulation of the scratch-slab addressing of the persistent and ranked node kernels
(mie_batch, ens_nodes_persist, ens_nodes_vjp_persist, ens_nodes_ranked / ens_nodes_vjp_ranked,
single_solve): every (wavefront, lane) of a launch touches only its own elements, all inside the
allocation, over every node it may take (any N up to the launch's rows - 2). The index
arithmetic mirrors the kernels' (scatrs_hip.hip mie_batch and mie_slab_per_wave;
ensemble_kernels.h slab_elems, sphere_ckpt_rows; mie_device.h ratio_chain_ckpt, x_chain,
real_column; layered_device.h layered_solve_xl), element by element in units of 8 bytes.

A persistent wavefront reuses its slab node after node: every element a node reads is written
by that node first (the checkpoint column for the blocks of orders 1..n, u_k(x) above floor(x),
every layered ratio column for k = 1..n+1), which the emulation checks too, so nothing one node
leaves behind reaches the next."""

from __future__ import annotations

import numpy as np
import pytest

K_ZBLOCK, K_SPHERE_BLOCK = 16, 8


def default_n_max(x):
    return int(np.ceil(x + 6.0 * np.cbrt(x) + 3.0))


def upward_orders(x, n):
    return min(int(np.floor(x)), n) if x >= 1.0 else 0


def sphere_elements(n, x, block, ck_base, ck_stride, ux_base, ux_stride):
    """(written, read) 8-byte elements of one sphere node: the checkpoints of ratio_chain_ckpt
    (complex: two elements each) and the u_k(x) column of x_chain and mie_orders."""
    written, read = set(), set()
    jmax = (n - 1) // block
    for j in range(jmax + 1):  # checkpoint j: u at index 2 + (j + 1) block, stored if <= cmax
        for h in (0, 1):
            written.add(ck_base + 2 * j * ck_stride + h)
    # regen(j) reads checkpoint j for every block j with 1 + j block <= n whose top is below the
    # start index (always, for the blocks below the last)
    for j in range(0, (n - 1) // block + 1):
        for h in (0, 1):
            read.add(ck_base + 2 * j * ck_stride + h)
    n_s = upward_orders(x, n)
    if n_s < n:
        for k in range(max(n_s + 1, 1), n + 2):
            written.add(ux_base + k * ux_stride)
        for k in range(n_s + 1, n + 1):  # mie_orders' ratio region reads u_k and u_{k+1}
            read.add(ux_base + k * ux_stride)
            read.add(ux_base + (k + 1) * ux_stride)
    return written, read


def layered_elements(L, n, slab_base, rows):
    """(written, read) elements of one layered node on its lane's slab (layered_solve_xl):
    2 L - 1 complex ratio columns of `rows` rows (column c at c rows 64 complex elements, element
    k at k 64) and the real u_k(x) column after them."""
    cs = rows * 64  # complex elements per column
    written, read = set(), set()
    for c in range(2 * L - 1):
        for k in range(1, n + 2):
            for h in (0, 1):
                written.add(slab_base + 2 * (c * cs + k * 64) + h)
        for k in range(1, n + 2):  # the outer step and the layer sweep read k and k + 1 <= n + 1
            for h in (0, 1):
                read.add(slab_base + 2 * (c * cs + k * 64) + h)
    return written, read


@pytest.mark.parametrize("block,waves_per_cu", [(K_SPHERE_BLOCK, 8), (K_SPHERE_BLOCK, 16)])
def test_sphere_batch_slabs_are_disjoint_and_inside(block, waves_per_cu):
    """mie_batch: checkpoints at ck_base + wave rows_ck 64 + lane (complex), u_k(x) at
    ux_base + wave rows 64 + lane (double), ux_base after every wavefront's checkpoints."""
    rng = np.random.default_rng(3)
    cus = 2
    x = np.exp(rng.uniform(np.log(0.1), np.log(120.0), 400))
    n = np.array([default_n_max(v) for v in x])
    nmax = int(n.max())
    rows, rows_ck = nmax + 2, nmax // block + 2
    waves = min(cus * waves_per_cu, (x.size + 63) // 64)
    waves = (waves + 3) // 4 * 4
    total = waves * 64 * (rows_ck * 2 + rows)  # 8-byte elements: the allocation
    owner = {}
    order = np.argsort(-n)
    for chunk in range((x.size + 63) // 64):
        wave = chunk % waves  # any wavefront may take any chunk
        for lane in range(64):
            idx = chunk * 64 + lane
            if idx >= x.size:
                continue
            i = order[idx]
            ck_base = 2 * (wave * rows_ck * 64 + lane)
            ux_base = 2 * waves * rows_ck * 64 + wave * rows * 64 + lane
            w, r = sphere_elements(int(n[i]), x[i], block, ck_base, 64, ux_base, 64)
            assert r <= w, (i, sorted(r - w)[:4])  # a node reads only what it wrote
            for e in w:
                assert 0 <= e < total
                assert owner.setdefault(e, (wave, lane)) == (wave, lane), e


@pytest.mark.parametrize("L", [1, 2, 3, 5, 8])
def test_node_kernel_slabs_are_disjoint_and_inside(L):
    """ens_nodes_persist / ens_nodes_vjp_persist (L = 1: slab_elems(1, rows) complex elements
    per wavefront, checkpoints then u_k(x)), ens_nodes_ranked / ens_nodes_vjp_ranked and
    single_solve (L > 1: slab_elems(L, rows) = rows 64 2 L), for a launch of 6 wavefronts whose
    rows come from its largest node, each wavefront taking several nodes in turn as a persistent
    loop would."""
    rng = np.random.default_rng(10 + L)
    waves = 6
    x = np.exp(rng.uniform(np.log(0.3), np.log(40.0), waves * 64 * 3))
    n = np.array([default_n_max(v) for v in x])
    rows = int(n.max()) + 2
    ckrows = (rows - 2) // K_ZBLOCK + 2
    per_wave = 64 * (ckrows + (rows + 1) // 2) if L == 1 else rows * 64 * 2 * L  # complex
    total = 2 * waves * per_wave
    owner = {}
    for g in range(x.size):
        wave, lane = (g // 64) % waves, g % 64
        base = 2 * (wave * per_wave + lane)  # slab = scratch + wave slab_elems + lane
        if L == 1:
            ux_base = 2 * (wave * per_wave + ckrows * 64) + lane  # real_column(slab, ckrows 64)
            w, r = sphere_elements(int(n[g]), x[g], K_ZBLOCK, base, 64, ux_base, 64)
        else:
            w, r = layered_elements(L, int(n[g]), base, rows)
            cs = rows * 64
            ux_base = 2 * (wave * per_wave + (2 * L - 1) * cs) + lane
            n_s = upward_orders(x[g], int(n[g]))
            for k in range(max(n_s + 1, 1), int(n[g]) + 2):
                w.add(ux_base + k * 64)
        assert r <= w
        for e in w:
            assert 0 <= e < total, (L, g, e, total)
            assert owner.setdefault(e, (wave, lane)) == (wave, lane), (L, g, e)


def test_a_two_column_stride_would_overlap_layered_slabs():
    """The emulation catches a slab stride that ignores the layer count: with the sphere
    kernels' two columns per wavefront, a 2-layer node's columns reach the next wavefront's."""
    L, rows, waves = 2, 40, 3
    owner = {}
    clash = False
    for wave in range(waves):
        for lane in range(64):
            base = 2 * (wave * rows * 64 * 2 + lane)  # two columns, not 2 L
            w, _ = layered_elements(L, rows - 2, base, rows)
            for e in w:
                if owner.setdefault(e, (wave, lane)) != (wave, lane):
                    clash = True
    assert clash
