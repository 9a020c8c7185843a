"""Seeded synthetic edge streams (numpy restatement, test helper).

The definition is shared with the product generator kernels
(risgraph_amd/csrc/rg_gen.h) and a GPU test checks the two agree bit for bit:

* ``draw(seed, i, k) = splitmix64(splitmix64(seed) ^ (64*i + k))``;
* R-MAT edge i at scale S: for level l = 0..S-1 (most significant bit first)
  ``r = draw(seed, i, l) >> 32`` picks quadrant a / b / c / d by the integer
  thresholds floor(a*2^32), floor((a+b)*2^32), floor((a+b+c)*2^32); quadrant b
  sets the dst bit, c the src bit, d both.  Self-loops and duplicates are
  kept.  Edges are i.i.d., so index order is already a uniformly random order
  (no separate shuffle pass);
* Erdos-Renyi edge i: ``src = draw(seed,i,0) % V``, ``dst = draw(seed,i,1) % V``;
* weight of stream edge i: ``1 + draw(wseed, i, 0) % 1000``.
"""
import numpy as np

M64 = (1 << 64) - 1
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
C1 = np.uint64(0xBF58476D1CE4E5B9)
C2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * C1
        z = (z ^ (z >> np.uint64(27))) * C2
    return z ^ (z >> np.uint64(31))


def draw(seed, i, k):
    key = splitmix64(np.uint64(seed))
    with np.errstate(over="ignore"):
        ctr = np.asarray(i, dtype=np.uint64) * np.uint64(64) + np.uint64(k)
    return splitmix64(key ^ ctr)


def thresholds(a, b, c):
    return int(a * 4294967296.0), int((a + b) * 4294967296.0), int((a + b + c) * 4294967296.0)


def rmat(scale, n, seed=1, a=0.57, b=0.19, c=0.19, start=0):
    ta, tab, tabc = (np.uint64(t) for t in thresholds(a, b, c))
    i = np.arange(start, start + n, dtype=np.uint64)
    src = np.zeros(n, np.uint64)
    dst = np.zeros(n, np.uint64)
    one = np.uint64(1)
    for lvl in range(scale):
        r = draw(seed, i, lvl) >> np.uint64(32)
        sb = (r >= tab).astype(np.uint64)
        db = (((r >= ta) & (r < tab)) | (r >= tabc)).astype(np.uint64)
        src = (src << one) | sb
        dst = (dst << one) | db
    return src, dst


def erdos_renyi(V, n, seed=1, start=0):
    i = np.arange(start, start + n, dtype=np.uint64)
    return draw(seed, i, 0) % np.uint64(V), draw(seed, i, 1) % np.uint64(V)


def weights(n, seed=2, start=0):
    i = np.arange(start, start + n, dtype=np.uint64)
    return np.uint64(1) + draw(seed, i, 0) % np.uint64(1000)


def max_out_degree_root(src, V):
    deg = np.bincount(src.astype(np.int64), minlength=V)
    return int(np.argmax(deg))          # ties -> lowest id
