"""The oracle's contact-aware dynamic palette against an independent
restatement of the reference's coloring (40_solver.metal:245-436), CPU only.

The restatement below works on plain neighbour sets built from the scene's
joints and springs and from the frame's active manifolds, and follows the
reference's rules: 20 synchronous passes, in which a body whose colour
equals a smaller-index dynamic neighbour's takes the lowest colour unused by
its smaller-index neighbours (capped at 63) and any other body moves down to
the lowest colour unused by all its dynamic neighbours when that is lower;
a pass after a pass that changed nothing changes nothing; when the last pass
still changed a colour, one serial greedy sweep in index order (lowest colour
unused by the already-swept smaller-index neighbours) finishes the palette.
Colours carry over from frame to frame and start at index mod 64
(GPUSolver.swift:967).  The palette is compared every frame, so a dropped
compaction, a wrong tie rule or a repair run at the wrong time fails."""
import numpy as np
import pytest

import oracle
from avbd_amd import demos

MAX_COLORS = 64


def _lowest_free(used):
    for c in range(MAX_COLORS):
        if c not in used:
            return c
    return None


def restated_palette(dyn, nbrs, prev):
    """One frame of the dynamic coloring; returns (colors, repaired)."""
    n = len(dyn)
    src = list(prev)
    last_changed = True
    for p in range(20):
        if p > 0 and not last_changed:
            continue
        dst = [0] * n
        changed = False
        for v in range(n):
            if not dyn[v]:
                continue
            lower = {src[u] for u in nbrs[v] if dyn[u] and u < v}
            every = {src[u] for u in nbrs[v] if dyn[u]}
            if src[v] in lower:
                c = _lowest_free(lower)
                dst[v] = MAX_COLORS - 1 if c is None else c
                changed = True
            else:
                c = _lowest_free(every)
                if c is not None and c < src[v]:
                    dst[v] = c
                    changed = True
                else:
                    dst[v] = src[v]
        src = dst
        last_changed = changed
    if not last_changed:
        return src, False
    for v in range(n):
        if not dyn[v]:
            src[v] = 0
            continue
        c = _lowest_free({src[u] for u in nbrs[v] if dyn[u] and u < v})
        src[v] = MAX_COLORS - 1 if c is None else c
    return src, True


def _static_neighbours(scene):
    n = len(scene.bodies)
    nbrs = [set() for _ in range(n)]
    for j in scene.joints:
        a, b = j.bodyA, j.bodyB
        if a >= 0 and b >= 0:
            nbrs[a].add(b)
            nbrs[b].add(a)
    for sp in scene.springs:
        a, b = sp["a"], sp["b"]
        if a >= 0 and b >= 0:
            nbrs[a].add(b)
            nbrs[b].add(a)
    return nbrs


def _long_chain(links=100):
    """A chain of boxes hinged in index order: neighbouring links start on
    neighbouring colours, so the 20 synchronous passes leave it unfinished
    and the serial repair runs."""
    s = demos.chain_scene(links=links, seed=9)
    return s


SCENES = {"batched_24": lambda: demos.batched_scene(24)[0], "stack_40": lambda: demos.stack_scene(height=40),
          "chain_100": _long_chain}


@pytest.mark.parametrize("name", list(SCENES))
def test_dynamic_palette_matches_restatement(name):
    s = SCENES[name]()
    o = oracle.OracleSolver(s)
    dyn = [(b["mass"] if b["mass"] is not None else b["density"]) > 0 for b in s.bodies]
    static = _static_neighbours(s)
    prev = [i % MAX_COLORS for i in range(len(s.bodies))]
    repairs = 0
    for frame in range(40):
        o.step()
        h, _, _ = o.manifolds()
        nbrs = [set(x) for x in static]
        for a, b, cnt in h:
            if cnt > 0:
                nbrs[int(a)].add(int(b))
                nbrs[int(b)].add(int(a))
        want, repaired = restated_palette(dyn, nbrs, prev)
        got = [int(c) for c in o.debugColors()]
        assert got == want, (name, frame)
        repairs += repaired
        prev = got
    if name == "chain_100":
        assert repairs > 0          # the serial repair path was exercised
