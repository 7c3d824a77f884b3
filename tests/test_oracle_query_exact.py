"""The oracle's query side against an exact-rational brute force (CPU).

raytrace (VDBMapping.hpp:675-721, intersector VDBMapping.hpp:1436-1449) and the
fast-mode cast (castRayIntoGridFast, VDBMapping.hpp:577-602) of the oracle walk a
4096^3 / 128^3 / 8^3 cell hierarchy in double precision. tests/rt_exact.py
restates both as a flat march over 8^3 leaf cells and a voxel walk in exact
rational arithmetic. Random small maps at resolution 1 hold active leaves, and
leaves with inactive voxels only, some of them outside the active box; the rays
are random, start inside active voxels, aim at active voxels, or cross a leaf
outside the box before entering it. Rays near a tie are left out; the rest must
agree. The last test plants mistakes in copies of the oracle (clip skipped, the
first voxel of the run tested, the box pad one voxel off either way, a later run
than the first, the open span dropped, only the first span walked) and checks
that the brute force catches every one.
"""
import math

import numpy as np
import pytest

import oracle as O
import rt_exact as R

SEEDS = [11, 12, 13, 14, 15]


def oracle_scene(seed, L=None):
    add_v, rem_v = R.random_scene(seed)
    m = O.OracleMap(1.0, L=L)
    assert m.set_config(O.make_config(max_range=500.0, fast_mode=True)) == O.OK
    assert m.add_points_to_grid(add_v + 0.5) == O.OK
    assert m.remove_points_from_grid(rem_v + 0.5) == O.OK
    assert m.integrate() == O.OK
    return m, R.scene_sets(add_v, rem_v)


def raytrace_mismatches(m, leaves, active, rays):
    """(rays compared, rays with a leaf run, indices that disagree)."""
    org, dirs, lens = rays
    st, ok, ends = m.raytrace(org, dirs, lens)
    assert st == O.OK
    clean = runs = 0
    bad = []
    for i in range(len(org)):
        E, D = R.exact_ray(org[i], dirs[i], lens[i])
        try:
            want = R.raytrace(leaves, active, E, D)
        except R.Unclean:
            continue
        clean += 1
        if want is None:
            seg_end = np.array([float(e + d) for e, d in zip(E, D)])
            if ok[i] or not np.allclose(ends[i], seg_end, rtol=0, atol=1e-9):
                bad.append(i)
        else:
            runs += 1
            v = tuple(int(x) for x in np.round(ends[i]))
            if not ok[i] or not np.array_equal(ends[i], np.round(ends[i])) or v not in want:
                bad.append(i)
    return clean, runs, bad


def fast_mismatches(m, active, segs):
    """(segments compared, segments marking a voxel, indices that disagree)."""
    S, E = segs
    clean = marking = 0
    bad = []
    for i in range(len(S)):
        ln = R.fast_clip_length(active, S[i], E[i])
        if ln is not None and 0 <= ln < 2 * R.RAY_EPS:
            continue  # grazes the box: the Ray::valid check of hits is not pinned here
        try:
            want = R.fast_marks(active, S[i], E[i])
        except R.Unclean:
            continue
        clean += 1
        marking += bool(want)
        got = {tuple(int(c) for c in v) for v in m.cast_ray_fast(S[i], E[i])}
        if got != want:
            bad.append(i)
    return clean, marking, bad


def test_walk_closed_forms():
    """The exact walk itself: an axis ray from a voxel centre visits 1 + |d| voxels;
    along (4, 1, 0) the y face is crossed at t = 1/2, between the x faces at 3/8 and
    5/8; (1, 1, 0) crosses x and y together and is refused; a start on a face of a
    negative-direction axis lies in the voxel below the face."""
    F = R.Fr
    cells, times, on_face, beyond, e_face = R.walk([F(1, 2)] * 3, [F(5), F(0), F(0)], F(0), F(1), 1,
                                                   extra=2)
    assert cells == [(i, 0, 0) for i in range(6)] and not on_face and not e_face
    assert beyond == [(6, 0, 0), (7, 0, 0)]
    cells, times, _, _, _ = R.walk([F(1, 2)] * 3, [F(4), F(1), F(0)], F(0), F(1), 1)
    assert cells == [(0, 0, 0), (1, 0, 0), (2, 0, 0), (2, 1, 0), (3, 1, 0), (4, 1, 0)]
    assert times == [0, F(1, 8), F(3, 8), F(1, 2), F(5, 8), F(7, 8)]
    with pytest.raises(R.Unclean):
        R.walk([F(1, 2)] * 3, [F(1), F(1), F(0)], F(0), F(1), 1)
    cells, _, on_face, _, _ = R.walk([F(8), F(1, 3), F(1, 3)], [F(-2), F(0), F(0)], F(0), F(1), 1)
    assert on_face and cells == [(7, 0, 0), (6, 0, 0)]


def wall_scene():
    """The hand-built map of the oracle's raytrace tests at resolution 1: a 9x9 wall
    of active voxels at x = 10 seen from the origin (the rays leave miss voxels in
    the leaves at x = 0 and x = 8)."""
    ys, zs = np.meshgrid(np.arange(-4, 5), np.arange(-4, 5))
    active = {(10, int(y), int(z)) for y, z in zip(ys.ravel(), zs.ravel())}
    leaves = {(x, y, z) for x in (0, 8) for y in (-8, 0) for z in (-8, 0)}
    return leaves, active


def test_brute_force_matches_hand_cases():
    """The brute force gives the hand-derived answers the oracle tests hold: the wall
    hit at voxel 10 from the origin voxel, and a run of one leaf cell without an
    active voxel (toward -x, length 4) reported past the run: voxel -1 or -2,
    depending on whether the crossing that ends the run counts as inside it."""
    leaves, active = wall_scene()
    F = R.Fr
    E = [F(1, 2)] * 3
    assert R.raytrace(leaves, active, E, [F(20), F(0), F(0)]) == {(10, 0, 0)}
    assert R.raytrace(leaves, active, E, [F(-4), F(0), F(0)]) == {(-1, 0, 0), (-2, 0, 0)}
    assert R.raytrace(leaves, active, [F(1, 2), F(61, 2), F(1, 2)], [F(0), F(4), F(0)]) is None


@pytest.mark.parametrize("seed", SEEDS)
def test_raytrace_matches_exact_brute_force(seed):
    m, (leaves, active) = oracle_scene(seed)
    rays = R.raytrace_rays(seed, leaves, active)
    clean, runs, bad = raytrace_mismatches(m, leaves, active, rays)
    assert not bad, f"{len(bad)} of {clean} rays disagree, first {bad[:5]}"
    assert clean >= 0.9 * len(rays[0]) and runs >= 0.4 * clean


@pytest.mark.parametrize("seed", SEEDS)
def test_fast_cast_matches_exact_brute_force(seed):
    m, (leaves, active) = oracle_scene(seed)
    segs = R.fast_rays(seed, active)
    clean, marking, bad = fast_mismatches(m, active, segs)
    assert not bad, f"{len(bad)} of {clean} segments disagree, first {bad[:5]}"
    # integer segments tie whenever two axes' extents over their gcd are both odd
    assert clean >= 0.2 * len(segs[0]) and marking >= 0.3 * clean


RT_NO_CLIP = ("    clip_ray(is, ray);  // setIndexRay: the result is ignored (VM:698) [U6]\n", "")
RT_FIRST_VOXEL = ("      while (dd.step() && !map_is_on(m, dd.vox[0], dd.vox[1], dd.vox[2])) {",
                  "      while (!map_is_on(m, dd.vox[0], dd.vox[1], dd.vox[2]) && dd.step()) {")
NO_PAD = ("  for (int a = 0; a < 3; ++a) is->bmax[a] += 1;  // mBBox.max().offset(1)", "")
RT_LATER_RUN = ("""      if (t.t0 < 0) t.t0 = d.time();
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) return true;
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
  return false;
}

bool hdda_128(""", """      if (t.t0 < 0) t.t0 = d.time();
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (!t.valid()) t.t0 = t.t1 = -1;
      else { Span keep = t; (void)keep; t.t0 = t.t1 = -1; if (!d.step()) { t = keep; return true; } continue; }
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
  return false;
}

bool hdda_128(""")
FAST_DROP_OPEN = ("  if (t.valid()) spans.push_back(t);\n  for (const Span& sp : spans) {",
                  "  for (const Span& sp : spans) {")
FAST_FIRST_SPAN = ("      if (map_is_on(m, d.vox[0], d.vox[1], d.vox[2])) visit(d.vox[0], d.vox[1], d.vox[2]);\n"
                   "    } while (d.step());\n",
                   "      if (map_is_on(m, d.vox[0], d.vox[1], d.vox[2])) visit(d.vox[0], d.vox[1], d.vox[2]);\n"
                   "    } while (d.step());\n    break;\n")

MUTANTS = {
    "raytrace: clip skipped": ("rt", [RT_NO_CLIP], ()),
    "raytrace: first voxel of the run tested": ("rt", [RT_FIRST_VOXEL], ()),
    "raytrace: box padded one voxel more": ("rt", [], ("ORC_U6_CLIP_PAD=1",)),
    "raytrace: box not padded": ("rt", [NO_PAD], ()),
    "raytrace: a later run than the first": ("rt", [RT_LATER_RUN], ()),
    "fast cast: box not padded": ("fast", [NO_PAD], ()),
    "fast cast: open span dropped": ("fast", [FAST_DROP_OPEN], ()),
    "fast cast: first span only": ("fast", [FAST_FIRST_SPAN], ()),
}


@pytest.mark.parametrize("name", list(MUTANTS))
def test_brute_force_catches_planted_mistake(name, tmp_path_factory):
    kind, repl, defines = MUTANTS[name]
    L = O.mutant(repl, defines, workdir=str(tmp_path_factory.getbasetemp() / "orc_mutants"))
    caught = 0
    for seed in SEEDS[:2]:
        m, (leaves, active) = oracle_scene(seed, L=L)
        if kind == "rt":
            _, _, bad = raytrace_mismatches(m, leaves, active, R.raytrace_rays(seed, leaves, active))
        else:
            _, _, bad = fast_mismatches(m, active, R.fast_rays(seed, active))
        caught += len(bad)
    assert caught > 0, f"planted mistake not caught: {name}"
