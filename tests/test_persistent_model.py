"""A host model of the persistent sweep's work queue (launch mode 2, switched
off), on the CPU: the tile list the library builds (build_tiles in
allium_amd/csrc/sweep_api.cpp, restated here) and the ticket/wait protocol of
k_sweep_persistent (allium_amd/csrc/sweep_kernels.hip), for the seven parity
cases of tests/test_gpu_sweep.py.

It checks what the protocol's termination rests on: tiles are handed out in
(level, direction) order, every awaited key's tiles precede the waiter in
ticket order, the count a waiter expects equals what the key's tiles add,
and completing key (L - 1, a) implies every upwind cell of key (L, a) is done.
A discrete-event replay with the grid's workgroups then drains every case.
Finally it states the duration the kernel's bounded waits allow when a
waiter never sees the counters move: one spin limit per dependent level."""
import numpy as np
import pytest

import allium_amd as A
from allium_amd import problems as P

# (name, mesh, S_N, jitter, lognormal sigma drawn first from the seed, as the parity test does)
CASES = [("2d_cfg1", (32, 32), 8, None, False), ("2d_p2_jitter", (12, 10), 6, 0.2, False),
         ("2d_p0", (16, 16), 4, None, False), ("3d_hetero", (12, 12, 12), 8, None, True),
         ("3d_jitter", (8, 7, 6), 6, 0.2, True), ("3d_p0", (10, 10, 10), 4, None, False),
         ("3d_reflect", (6, 6, 6), 4, None, False)]
TILE = 256          # persistent_tile: threads per workgroup = cells per tile
SPIN_LIMIT = 1 << 21


def _plan(name, nel, sn, jitter, lognormal):
    rng = np.random.default_rng(17)
    if lognormal:
        P.lognormal_sigma(int(np.prod(nel)), 1.0, rng)
    v = P.jittered_vertices(nel, jitter, rng) if jitter else None
    bc = (1, 0, 0, 0, 0, 0) if name == "3d_reflect" else None
    m = A.Mesh(nel, vertices=v, bc=bc)
    lev, nl = A.host_levels(m, sn)
    return m, lev, nl, bc is None


def _tiles(lev, nl, angle_independent, G=1):
    """build_tiles: cells grouped by level, then direction, then element."""
    Ne, No = lev.shape
    e, a = np.meshgrid(np.arange(Ne), np.arange(No), indexing="ij")
    order = np.lexsort((e.ravel(), a.ravel(), lev.ravel()))
    L, D = lev.ravel()[order], a.ravel()[order]
    tiles, need = [], {}
    keyf = (lambda l, d: l * No + d) if angle_independent else (lambda l, d: l)
    prevf = (lambda l, d: (l - 1) * No + d) if angle_independent else (lambda l, d: l - 1)
    i = 0
    while i < len(order):
        j = i
        k = keyf(L[i], D[i])
        while j < len(order) and keyf(L[j], D[j]) == k:
            j += 1
        nc = j - i
        need[k] = nc * G
        for it in range(0, nc * G, TILE):
            tiles.append((k, prevf(L[i], D[i]) if L[i] > 0 else -1, min(TILE, nc * G - it)))
        i = j
    return tiles, need


@pytest.mark.parametrize("name,nel,sn,jitter,lognormal", CASES, ids=[c[0] for c in CASES])
def test_tile_queue_invariants(name, nel, sn, jitter, lognormal):
    m, lev, nl, indep = _plan(name, nel, sn, jitter, lognormal)
    tiles, need = _tiles(lev, nl, indep)
    first, last = {}, {}
    for t, (k, w, n) in enumerate(tiles):
        first.setdefault(k, t)
        last[k] = t
    # need = what the key's tiles add
    added = {}
    for k, w, n in tiles:
        added[k] = added.get(k, 0) + n
    assert added == need
    for t, (k, w, n) in enumerate(tiles):
        if w < 0:
            continue
        assert w in need and need[w] > 0        # never waits on an empty key
        assert last[w] < t                       # the awaited tiles were handed out earlier
    # tickets in nondecreasing level order
    No = lev.shape[1]
    lv = [k // No if indep else k for k, _, _ in tiles]
    assert all(x <= y for x, y in zip(lv, lv[1:]))
    # per direction the levels are contiguous from 0, so (L - 1, a) done implies all of (< L, a) done
    if indep:
        for d in range(No):
            ls = np.unique(lev[:, d])
            assert np.array_equal(ls, np.arange(ls.size))


def _replay(tiles, need, resident):
    """Workgroups take tickets in order, wait until the awaited key is
    complete, work one step and add their items.  Returns the number of steps
    to drain, or None if the queue stalls."""
    done = {k: 0 for k in need}
    nxt, t, active = 0, 0, []      # (tile index)
    while nxt < len(tiles) or active:
        while len(active) < resident and nxt < len(tiles):
            active.append(nxt)
            nxt += 1
        ready = [i for i in active if tiles[i][1] < 0 or done[tiles[i][1]] >= need[tiles[i][1]]]
        if not ready:
            return None
        for i in ready:
            done[tiles[i][0]] += tiles[i][2]
            active.remove(i)
        t += 1
    return t


@pytest.mark.parametrize("name,nel,sn,jitter,lognormal", CASES, ids=[c[0] for c in CASES])
def test_queue_drains_with_coherent_counters(name, nel, sn, jitter, lognormal):
    """With counters every workgroup sees, the grid drains: the replay takes
    about one step per dependent level whatever the number of resident
    workgroups (256 CUs x 1..4)."""
    m, lev, nl, indep = _plan(name, nel, sn, jitter, lognormal)
    tiles, need = _tiles(lev, nl, indep)
    for resident in (256, 1024):
        steps = _replay(tiles, need, resident)
        assert steps is not None and steps <= len(tiles)
        assert steps >= nl


def test_bounded_waits_allow_minutes_when_counters_are_not_seen():
    """If a waiter's polling loads never observed other workgroups' counter
    updates (nor the error flag), every dependent level would run one wait to
    its spin limit in turn: for the first case (32^2 quads, S8: 63 levels),
    62 x 2^21 polls at ~1 microsecond per agent-scope load and s_sleep 1 is
    about 130 s, consistent with the run that did not end within 120 s."""
    m, lev, nl, indep = _plan(*CASES[0])
    assert nl == 63
    seconds = (nl - 1) * SPIN_LIMIT * 1e-6
    assert 120 < seconds < 140
