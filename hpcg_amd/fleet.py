"""Grid sizes of a fleet of independent problems (hpcgx_generate_fleet).

The fleet workload draws every grid axis of every problem from a fixed size
set with a splitmix64 stream, so any host reproduces the same fleet:
problem q's (nx, ny, nz) are the stream's outputs 3q, 3q+1, 3q+2, each mapped
to [0, 1) with its top 53 bits and scaled to an index into the size set.
"""
import numpy as np

FLEET_SIZES = (32, 48, 64, 96, 128)


def splitmix64(seed, n):
    """The first n outputs of splitmix64 seeded with `seed` (uint64)."""
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, n + 1, dtype=np.uint64)
        z = (s ^ (s >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def fleet_dims(count=32, seed=7, sizes=FLEET_SIZES):
    """[(nx, ny, nz)] of `count` problems drawn from `sizes` by splitmix64(seed)."""
    u = (splitmix64(seed, 3 * count) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    idx = np.minimum((u * len(sizes)).astype(np.int64), len(sizes) - 1)
    d = [int(sizes[k]) for k in idx]
    return [tuple(d[3 * q:3 * q + 3]) for q in range(count)]
