"""Exact-rational brute force of the query side: raytrace and the fast-mode cast.

What it restates, independently of the oracle's hierarchical double-precision
code:

- ``raytrace`` (VDBMapping.hpp:675-721, intersector from :1436-1449): the ray is
  clipped to the padded bounding box of the leaves that hold an active voxel
  (the box ends one voxel past the last voxel, i.e. on the far face of the last
  leaf cell), a clipped ray no longer than 1e-5 (float epsilon) finds nothing,
  otherwise the first maximal run of consecutive 8^3 leaf cells that hold a leaf
  is the march interval, and a voxel walk over that interval reports the first
  active voxel after its first voxel, or, when there is none, the voxel its last
  step (the first one past the interval's end) moved to. A flat march
  over leaf cells is used here: a hierarchy of 4096^3 / 128^3 / 8^3 cells visits
  the same leaf cells in the same order, so its first interval is the same run.
- the fast-mode cast (VDBMapping.hpp:577-602): every run of leaf cells is one span
  and the voxel walk over a span marks its active voxels; every active voxel lies
  in a leaf and inside the box, so the marks are exactly the active voxels on the
  voxel path from the centre of s to the centre of e.

All times and positions are ``fractions.Fraction``: crossing times are exact, so a
ray that crosses two faces of different axes at the same time is detected and
left out (double-precision DDAs break such ties by accumulated rounding), and so
is a ray with any two face crossings of different axes, or a crossing and an end
of the ray, closer than a margin in time. Where the reference's result depends
on how a double lands on a face (the walk starting exactly on the face where the
run or the box begins, and stopping exactly on the face where it ends), the
brute force returns every outcome that landing allows: the voxel before the
entry face may be the walk's (untested) first voxel, and the crossing of the exit
face may count as inside the interval or past it.
"""
from __future__ import annotations

import math
from fractions import Fraction as Fr

import numpy as np

RAY_EPS = Fr(float(np.float32(1e-5)))  # Ray::valid() threshold, float epsilon as a double


class Unclean(Exception):
    """The ray passes too close to a tie for an exact answer to be comparable."""


def walk(E, D, t0, t1, size, extra=0):
    """Exact walk of the segment E + t D, t0 <= t <= t1, through cubic cells of edge
    `size` (cell c spans [c*size, (c+1)*size) on each axis).

    Returns (cells, times, start_on_face, beyond, end_on_face): cells[i] is entered
    at times[i] < t1 (times[0] = t0, and cells[0] is the cell on the far side of any
    face E + t0 D lies on); start_on_face tells whether E + t0 D lies on a face of
    an axis the ray moves along; beyond holds the next `extra` cells the line
    enters at times >= t1; end_on_face tells whether the first of those crossings
    falls exactly on t1. Raises Unclean when two axes cross at the same time."""
    cur, nxt, step = [], [], []
    on_face = False
    for a in range(3):
        p = E[a] + t0 * D[a]
        q = p / size
        if D[a] > 0:
            c = math.floor(q)
            nt = (Fr((c + 1) * size) - E[a]) / D[a]
        elif D[a] < 0:
            c = math.ceil(q) - 1
            nt = (Fr(c * size) - E[a]) / D[a]
        else:
            c = math.floor(q)
            nt = None
        if D[a] != 0 and q.denominator == 1:
            on_face = True
        cur.append(c)
        nxt.append(nt)
        step.append(Fr(size) / abs(D[a]) if D[a] != 0 else None)
    cells, times, beyond = [tuple(cur)], [t0], []
    end_on_face = False
    while True:
        live = [(nxt[a], a) for a in range(3) if nxt[a] is not None]
        if not live or len(beyond) >= extra and min(t for t, _ in live) >= t1:
            return cells, times, on_face, beyond, end_on_face
        tmin = min(t for t, _ in live)
        axes = [a for t, a in live if t == tmin]
        if len(axes) > 1:
            raise Unclean("two axes cross at the same time")
        a = axes[0]
        cur[a] += 1 if D[a] > 0 else -1
        nxt[a] += step[a]
        if tmin >= t1:
            end_on_face |= not beyond and tmin == t1
            beyond.append(tuple(cur))
        else:
            cells.append(tuple(cur))
            times.append(tmin)


def check_margin(E, D, delta):
    """Raise Unclean if two unit-face crossings of different axes are closer than
    delta in time, or a crossing is within delta of t = 0 or 1; crossings up to
    a few voxels past t = 1 count, since a walk may stop past its end."""
    ev = []
    moving = [abs(float(d)) for d in D if d != 0]
    tmax = 1 + 3 / min(moving) if moving else 1  # two crossings past the end count too
    for a in range(3):
        if D[a] == 0:
            continue
        e, d = float(E[a]), float(D[a])
        lo, hi = sorted((e, e + d * tmax))
        for k in range(math.floor(lo) - 1, math.ceil(hi) + 2):
            t = (k - e) / d
            if -delta <= t <= tmax:
                if abs(t) < delta or abs(t - 1) < delta:
                    raise Unclean("crossing at an end of the ray")
                ev.append((t, a))
    ev.sort()
    for (ta, aa), (tb, ab) in zip(ev, ev[1:]):
        if aa != ab and tb - ta < delta:
            raise Unclean("near tie")


def padded_box(active):
    """The intersector's box: the node bounding box of every leaf with an active
    voxel, its max corner moved one voxel out; None if no voxel is active."""
    if not active:
        return None
    org = np.array([[(c >> 3) << 3 for c in v] for v in active], np.int64)
    return tuple(int(x) for x in org.min(0)), tuple(int(x) + 8 for x in org.max(0))


def clip(box, E, D, delta):
    """Ray::clip against the box: (t0, t1) of the clipped ray, or (0, 1) unchanged
    when the ray misses the box."""
    t0, t1 = Fr(0), Fr(1)
    if box is None:
        return t0, t1
    lo, hi = box
    a0, a1 = Fr(0), Fr(1)
    for a in range(3):
        if D[a] == 0:
            if lo[a] == E[a] or hi[a] == E[a]:
                raise Unclean("axis-parallel ray on a box face")
            if not lo[a] < E[a] < hi[a]:
                return t0, t1
            continue
        ta = (lo[a] - E[a]) / D[a]
        tb = (hi[a] - E[a]) / D[a]
        if ta > tb:
            ta, tb = tb, ta
        a0, a1 = max(a0, ta), min(a1, tb)
    if abs(a1 - a0) < delta:
        raise Unclean("ray grazes the box")
    return (a0, a1) if a0 <= a1 else (t0, t1)


def raytrace(leaves, active, E, D, delta=Fr(1, 10**7)):
    """Outcomes of raytrace for one index-space ray E + t D, t in [0, 1]:
    None when no leaf run is found (success false), else the set of voxels the
    reference may report as the end point."""
    check_margin(E, D, float(delta))
    t0, t1 = clip(padded_box(active), E, D, delta)
    if abs((t1 - t0) - RAY_EPS) < delta:
        raise Unclean("clipped length at the Ray::valid threshold")
    if not (t1 - t0 > RAY_EPS):
        return None
    cells, times, _, _, _ = walk(E, D, t0, t1, 8)
    start = end = None
    for c, t in zip(cells, times):
        if tuple(8 * x for x in c) in leaves:
            if start is None:
                start = t
        elif start is not None:
            end = t
            break
    if start is None:
        return None
    if end is None:
        end = t1
    # The walk tests the voxel after each step whose crossing time is <= end and
    # stops on the first active one; a step past end still moves to the next voxel,
    # which is then reported.
    vox, _, s_face, beyond, e_face = walk(E, D, start, end, 1, extra=2)
    k = len(vox) - 1
    seq = vox + beyond
    outcomes = set()
    for s in ((0, 1) if s_face else (1,)):
        for e in ((k, k + 1) if e_face else (k,)):
            found = next((seq[j] for j in range(s, e + 1) if seq[j] in active), None)
            outcomes.add(found if found is not None else seq[e + 1])
    return outcomes


def fast_marks(active, s, e):
    """Voxels the fast-mode cast from voxel s to voxel e marks: the active voxels on
    the exact voxel path between the two centres. Raises Unclean on a tie."""
    E = [Fr(int(c)) + Fr(1, 2) for c in s]
    D = [Fr(int(b) - int(a)) for a, b in zip(s, e)]
    if all(d == 0 for d in D):
        return {tuple(int(c) for c in s)} & active
    cells, _, _, _, _ = walk(E, D, Fr(0), Fr(1), 1)
    return {c for c in cells if c in active}


def fast_clip_length(active, s, e):
    """Length in t of the fast-mode ray (centre of s to centre of e) clipped to the
    box, for leaving out rays that only graze it."""
    E = [Fr(int(c)) + Fr(1, 2) for c in s]
    D = [Fr(int(b) - int(a)) for a, b in zip(s, e)]
    box = padded_box(active)
    if box is None:
        return None
    lo, hi = box
    a0, a1 = Fr(0), Fr(1)
    for a in range(3):
        if D[a] == 0:
            if not lo[a] <= E[a] <= hi[a]:
                return Fr(-1)
            continue
        ta, tb = sorted(((lo[a] - E[a]) / D[a], (hi[a] - E[a]) / D[a]))
        a0, a1 = max(a0, ta), min(a1, tb)
    return a1 - a0


# ---------------------------------------------------------------- scenes --

def random_scene(seed):
    """A small random map at resolution 1: active leaves near the centre with
    active voxels of mixed density (and some inactive voxels), plus leaves that
    hold only inactive voxels, several of them outside the active box. Returns
    (add_voxels, remove_voxels): the voxels to mark occupied, then free."""
    rng = np.random.default_rng(seed)
    cells = np.array([(x, y, z) for x in range(-3, 4) for y in range(-3, 4) for z in range(-3, 4)])
    inner = cells[(np.abs(cells) <= 2).all(1)]
    act = inner[rng.choice(len(inner), size=int(rng.integers(6, 22)), replace=False)]
    add, rem = [], []
    for c in act:
        dens = rng.choice([0.01, 0.04, 0.15, 0.5])
        m = rng.random(512) < dens
        if not m.any():
            m[rng.integers(512)] = True
        idx = np.nonzero(m)[0]
        loc = np.stack([idx >> 6, (idx >> 3) & 7, idx & 7], 1)
        add.append(c * 8 + loc)
        free = rng.integers(0, 8, size=(int(rng.integers(0, 4)), 3))
        rem.append(c * 8 + free)
    act_set = {tuple(c) for c in act}
    others = [c for c in cells if tuple(c) not in act_set]
    ina = [others[i] for i in rng.choice(len(others), size=int(rng.integers(15, 60)), replace=False)]
    for c in ina:
        rem.append(c * 8 + rng.integers(0, 8, size=(int(rng.integers(1, 4)), 3)))
    add_v = np.concatenate(add).astype(np.int64)
    rem_v = np.concatenate(rem).astype(np.int64)
    return add_v, rem_v


def scene_sets(add_v, rem_v):
    """(leaf origins, active voxels) after marking add_v occupied, then rem_v free."""
    active = {tuple(int(x) for x in v) for v in add_v}
    active -= {tuple(int(x) for x in v) for v in rem_v}
    leaves = {tuple(int(x) >> 3 << 3 for x in v) for v in np.concatenate([add_v, rem_v])}
    return leaves, active


def raytrace_rays(seed, leaves, active, n=700):
    """Query rays (origins, directions, lengths) in world = index coordinates:
    random rays across the scene, rays that start inside an active voxel, rays
    aimed at active voxels from outside, and rays that pass through a leaf lying
    outside the active box before they enter the box (half of them near its far
    faces)."""
    rng = np.random.default_rng(seed)
    act = np.array(sorted(active), np.float64)
    box = padded_box(active)
    lo, hi = np.array(box[0], float), np.array(box[1], float)
    outside = [np.array(l, float) for l in leaves
               if any(l[a] < lo[a] or l[a] >= hi[a] for a in range(3))]
    near_far = [l for l in outside if any(l[a] == hi[a] for a in range(3))]
    O_, D_, L_ = [], [], []

    def unit(v):
        return v / np.linalg.norm(v)

    for i in range(n):
        kind = i % 5
        if kind == 0:  # anywhere
            o = rng.uniform(-36, 36, 3)
            d = unit(rng.normal(size=3))
            ln = rng.uniform(4, 120)
        elif kind == 1:  # from inside an active voxel
            o = act[rng.integers(len(act))] + rng.uniform(0.05, 0.95, 3)
            d = unit(rng.normal(size=3))
            ln = rng.uniform(2, 60)
        elif kind == 2:  # toward an active voxel
            t = act[rng.integers(len(act))] + rng.uniform(0.05, 0.95, 3)
            o = rng.uniform(-36, 36, 3)
            d = unit(t - o)
            ln = np.linalg.norm(t - o) * rng.uniform(0.7, 1.6)
        else:  # through a leaf outside the box, then into the box
            pool = near_far if (kind == 3 and near_far) else (outside or [np.zeros(3)])
            leaf = pool[rng.integers(len(pool))]
            p = leaf + rng.uniform(0.05, 7.95, 3)
            t = rng.uniform(lo + 0.05, hi - 0.05)
            d = unit(t - p)
            o = p - d * rng.uniform(0.0, 12.0)
            ln = np.linalg.norm(t - o) + rng.uniform(0, 30)
        O_.append(o)
        D_.append(d)
        L_.append(ln)
    return np.array(O_), np.array(D_), np.array(L_)


def fast_rays(seed, active, n=1200):
    """Fast-mode cast segments (start voxel, end voxel), integer index coordinates."""
    rng = np.random.default_rng(seed)
    act = np.array(sorted(active), np.int64)
    S, Ev = [], []
    for i in range(n):
        kind = i % 3
        if kind == 0:
            s = rng.integers(-36, 36, 3)
            e = rng.integers(-36, 36, 3)
        elif kind == 1:
            s = act[rng.integers(len(act))]
            e = rng.integers(-40, 40, 3)
        else:
            s = rng.integers(-40, 40, 3)
            e = act[rng.integers(len(act))] + rng.integers(-3, 4, 3)
        S.append(s)
        Ev.append(e)
    return np.array(S, np.int32), np.array(Ev, np.int32)


def exact_ray(o, d, ln):
    """The index-space ray (E, D) of a query at resolution 1: direction scaled to
    the query length. The oracle normalises in double precision; the difference is
    a few ulps, far inside the comparison margin."""
    dd = np.asarray(d, np.float64)
    dd = dd / math.sqrt(float(dd @ dd)) * float(ln)
    return [Fr(float(x)) for x in o], [Fr(float(x)) for x in dd]
