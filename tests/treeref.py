"""Independent numpy statement of the octree index sets, used to check the
device tree at full size without per-box Python loops:

* Morton keys: x is the most significant bit of each octal digit, coarse
  levels first (octree.c:132-142);
* I4(b) (octree.c:497-543): boxes whose parent is the parent of b or one of
  its neighbours, minus b's own neighbourhood (|ds| <= 1 in every axis);
* the M2L pairs of a level: I4(b) restricted to the non-empty boxes.

Pinned against the oracle's I4 (itself pinned by brute force) in
tests/test_oracle_tree.py."""
import numpy as np


def morton_decode(keys, l):
    keys = np.asarray(keys, dtype=np.int64)
    ijk = np.zeros((keys.size, 3), dtype=np.int64)
    for j in range(l):
        dig = (keys >> (3 * j)) & 7
        ijk[:, 0] |= ((dig >> 2) & 1) << j
        ijk[:, 1] |= ((dig >> 1) & 1) << j
        ijk[:, 2] |= (dig & 1) << j
    return ijk


def morton_encode(ijk, l):
    ijk = np.asarray(ijk, dtype=np.int64)
    key = np.zeros(ijk.shape[0], dtype=np.int64)
    for j in range(l):
        dig = (((ijk[:, 0] >> j) & 1) << 2) | (((ijk[:, 1] >> j) & 1) << 1) | ((ijk[:, 2] >> j) & 1)
        key |= dig << (3 * j)
    return key


OFFSETS = np.array([(dx, dy, dz) for dx in range(-3, 4) for dy in range(-3, 4)
                    for dz in range(-3, 4)], dtype=np.int64)


def i4_keys(key, l):
    """I4 of one box as a sorted key array (all boxes of the level grid)."""
    b = morton_decode([key], l)[0]
    s = b[None, :] - OFFSETS
    ok = np.all((s >= 0) & (s < (1 << l)), axis=1)
    ok &= np.all(np.abs((s >> 1) - (b >> 1)[None, :]) <= 1, axis=1)
    ok &= ~np.all(np.abs(s - b[None, :]) <= 1, axis=1)
    return np.sort(morton_encode(s[ok], l))


def m2l_pairs(keys, l):
    """(target key, source key, offset code) of every M2L pair among the
    non-empty boxes `keys` (sorted) of level l; offset = target - source in
    grid steps, code (dx+3)*49 + (dy+3)*7 + (dz+3)."""
    keys = np.asarray(keys, dtype=np.int64)
    b = morton_decode(keys, l)
    tg, sr, oc = [], [], []
    for c, d in enumerate(OFFSETS):
        s = b - d[None, :]
        ok = np.all((s >= 0) & (s < (1 << l)), axis=1)
        ok &= np.all(np.abs((s >> 1) - (b >> 1)) <= 1, axis=1)
        ok &= not np.all(np.abs(d) <= 1)
        if not ok.any():
            continue
        sk = morton_encode(s[ok], l)
        pos = np.searchsorted(keys, sk)
        hit = pos < keys.size
        hit[hit] = keys[pos[hit]] == sk[hit]
        tg.append(keys[ok][hit])
        sr.append(sk[hit])
        oc.append(np.full(int(hit.sum()), c, dtype=np.int64))
    if not tg:
        return (np.zeros(0, np.int64),) * 3
    return np.concatenate(tg), np.concatenate(sr), np.concatenate(oc)


def pair_codes(tgt_keys, src_keys):
    return np.asarray(tgt_keys, dtype=np.int64) * (1 << 32) + np.asarray(src_keys, dtype=np.int64)
