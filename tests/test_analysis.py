"""Cu-cluster statistics (akmc_cu_clusters) against a breadth-first search
written from BCC geometry (CPU only).

The search uses Cartesian coordinates in units of a0/2: corner sites at even
(X, Y, Z), body sites at odd ones; two sites are first neighbours when every
coordinate differs by exactly 1 (periodically).  It shares no code or index
convention with the library."""
from collections import deque

import numpy as np
import pytest

import akmc

V, CU, MN, NI = 1, 2, 4, 8


def site_xyz(i, box):
    nx, ny, nz = box
    X = 2 * nx
    x, y, z = i % X, (i // X) % ny, i // (X * ny)
    # doubled-x index: cell = x // 2, body = x & 1
    return (2 * (x // 2) + (x & 1), 2 * y + (x & 1), 2 * z + (x & 1))


def bfs_stats(sites, box):
    nx, ny, nz = box
    pos = {}
    for i in range(len(sites)):
        pos[site_xyz(i, box)] = i
    reads_cu = lambda f: (f & V) == 0 and (f & CU) != 0
    L = (2 * nx, 2 * ny, 2 * nz)

    def nbrs(p):
        for dx in (-1, 1):
            for dy in (-1, 1):
                for dz in (-1, 1):
                    yield pos[((p[0] + dx) % L[0], (p[1] + dy) % L[1], (p[2] + dz) % L[2])]

    seen = set()
    sizes = []
    isolated = 0
    for i, f in enumerate(sites):
        p = site_xyz(i, box)
        if f & CU and not any(reads_cu(sites[j]) for j in nbrs(p)):
            isolated += 1
        if not reads_cu(f) or i in seen:
            continue
        seen.add(i)
        q = deque([i])
        n = 0
        while q:
            a = q.popleft()
            n += 1
            for j in nbrs(site_xyz(a, box)):
                if j not in seen and reads_cu(sites[j]):
                    seen.add(j)
                    q.append(j)
        sizes.append(n)
    return sorted(sizes), isolated


def _sizes_from_hist(h):
    out = []
    for k, c in enumerate(h):
        out += [k] * int(c)
    return sorted(out)


@pytest.mark.parametrize("box,seed,frac", [((4, 4, 4), 1, 0.2), ((5, 4, 6), 2, 0.35), ((6, 6, 6), 3, 0.1),
                                           ((3, 5, 4), 4, 0.5)])
def test_clusters_match_bfs(box, seed, frac):
    rng = np.random.default_rng(seed)
    n = 2 * box[0] * box[1] * box[2]
    sites = np.where(rng.random(n) < frac, CU, 0).astype(np.uint8)
    # a few vacancies and stale tags: V|Cu reads as V, Mn|Cu as Cu
    sites[rng.integers(0, n, 5)] |= V
    sites[rng.integers(0, n, 5)] |= MN
    st = akmc.cu_clusters(sites, box, hist_cap=n + 1)
    sizes, isolated = bfs_stats(sites, box)
    assert _sizes_from_hist(st["hist"]) == sizes
    assert st["clusters"] == len(sizes) and st["cu"] == sum(sizes)
    assert st["isolated"] == isolated


def test_known_configurations():
    box = (4, 4, 4)
    X = 8
    idx = lambda x, y, z: x + X * (y + 4 * z)
    s = np.zeros(2 * 64, dtype=np.uint8)
    s[idx(0, 0, 0)] = CU                     # corner site ...
    s[idx(7, 3, 3)] = CU                     # ... and its body-centre neighbour across the periodic corner
    s[idx(4, 2, 2)] = CU                     # a lone atom
    st = akmc.cu_clusters(s, box)
    assert st["clusters"] == 2 and st["cu"] == 3 and st["isolated"] == 1
    assert st["hist"][1] == 1 and st["hist"][2] == 1
    s[idx(4, 2, 2)] |= V                     # a vacancy with a Cu tag is not an atom, but is in the Cu set
    st = akmc.cu_clusters(s, box)
    assert st["clusters"] == 1 and st["cu"] == 2 and st["isolated"] == 1
