"""The fleet workload's grid sizes (hpcg_amd.fleet_dims), on the CPU.

splitmix64 is checked against a scalar restatement of its published
recurrence (Steele, Lea and Flood, "Fast splittable pseudorandom number
generators", 2014) with Python integers, and the size draw against the
mapping it documents."""
from hpcg_amd import fleet as F

MASK = (1 << 64) - 1


def splitmix_scalar(seed, n):
    out, s = [], seed
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & MASK
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_matches_scalar_recurrence():
    for seed in (0, 7, 42, 2**63 + 5):
        assert [int(v) for v in F.splitmix64(seed, 50)] == splitmix_scalar(seed, 50)


def test_splitmix64_known_value():
    # the first output of splitmix64 seeded with 0
    assert int(F.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF


def test_fleet_dims_mapping():
    sizes = F.FLEET_SIZES
    raw = splitmix_scalar(7, 96)
    want = [sizes[min((v >> 11) * len(sizes) // (1 << 53), len(sizes) - 1)] for v in raw]
    dims = F.fleet_dims(32, 7)
    assert len(dims) == 32
    assert [a for d in dims for a in d] == want
    assert all(a in sizes for d in dims for a in d)
    # every size occurs; the fleet is uneven
    assert set(a for d in dims for a in d) == set(sizes)
    assert len(set(dims)) > 20
