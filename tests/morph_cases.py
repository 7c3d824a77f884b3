"""Leaf sets for the morphology tests (CPU pinning and GPU parity).

A leaf set is (origins (n, 3) int32, masks (n, 8) uint64) in OpenVDB's leaf
layout: voxel offset (x & 7) << 6 | (y & 7) << 3 | (z & 7), bit offset & 63
of word offset >> 6.
"""
import numpy as np

OPS = ("dilate", "erode", "close", "open")


def random_leafset(seed, span=4, keep=0.6, density=0.3, blobs=True, lo=-16):
    """Leaves on a span^3 grid of leaf slots starting at voxel `lo` (a
    multiple of 8), each kept with probability `keep`. Masks are random
    voxels at `density`, or, with blobs, random boxes, so that erosion leaves
    something behind."""
    rng = np.random.default_rng(seed)
    origins, masks = [], []
    for i in range(span):
        for j in range(span):
            for k in range(span):
                if rng.random() >= keep:
                    continue
                vox = np.zeros((8, 8, 8), bool)
                if blobs:
                    for _ in range(rng.integers(0, 4)):
                        a = rng.integers(0, 8, 3)
                        b = np.minimum(a + rng.integers(1, 9, 3), 8)
                        vox[a[0]:b[0], a[1]:b[1], a[2]:b[2]] = True
                    vox |= rng.random((8, 8, 8)) < density * 0.1
                else:
                    vox = rng.random((8, 8, 8)) < density
                origins.append([lo + 8 * i, lo + 8 * j, lo + 8 * k])
                masks.append(pack(vox))
    order = rng.permutation(len(origins))
    o = np.array(origins, np.int32).reshape(-1, 3)[order]
    m = np.array(masks, np.uint64).reshape(-1, 8)[order]
    return np.ascontiguousarray(o), np.ascontiguousarray(m)


def pack(vox):
    """(8, 8, 8) bool [x, y, z] -> 8 uint64 words."""
    flat = np.asarray(vox, bool).reshape(512)
    words = np.zeros(8, np.uint64)
    for n in np.flatnonzero(flat):
        words[n >> 6] |= np.uint64(1) << np.uint64(n & 63)
    return words


def unpack(words):
    """8 uint64 words -> (8, 8, 8) bool [x, y, z]."""
    w = np.asarray(words, np.uint64)
    bits = (w[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)
    return bits.astype(bool).reshape(512).reshape(8, 8, 8)


def to_dense(origins, masks, lo, size):
    """Scatter a leaf set into a dense bool array covering [lo, lo + size)^3."""
    d = np.zeros((size, size, size), bool)
    for o, m in zip(np.asarray(origins).reshape(-1, 3), np.asarray(masks).reshape(-1, 8)):
        x, y, z = (np.asarray(o) - lo)
        d[x:x + 8, y:y + 8, z:z + 8] |= unpack(m)
    return d
