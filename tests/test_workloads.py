"""The synthetic-input generators (splitmix64-seeded xoshiro256**): published
known answers of both generators, the seeding composition, agreement with the
oracle's generators, and the distributions' own properties.  CPU only."""
import numpy as np
import pytest

from acc_hip import workloads as W
import oracle_py as O


def test_splitmix64_known_answers():
    # splitmix64 from state 0 (Steele, Lea and Flood; the values quoted with
    # Vigna's reference implementation)
    assert W.splitmix64(0, 3).tolist() == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_xoshiro256ss_known_answers():
    # xoshiro256** from the state (1, 2, 3, 4): the reference implementation's
    # first outputs (the first is rotl(2 * 5, 7) * 9 = 11520 by hand)
    assert W.xoshiro256ss([1, 2, 3, 4], 4).tolist() == [11520, 0, 1509978240, 1215971899390074240]


def test_seeding_is_splitmix_then_xoshiro():
    # uniform_i32 is lo + (raw mod span) of the stream seeded by splitmix64(seed)
    for seed in (0, 42, 2**63 + 5):
        raw = W.xoshiro256ss(W.splitmix64(seed, 4), 1000)
        want = (-1000 + (raw % np.uint64(2001)).astype(np.int64)).astype(np.int32)
        assert np.array_equal(W.uniform_i32(seed, -1000, 1000, 1000), want)
        # the unit doubles behind the normals: u1 = 1 - top53 * 2^-53 lies in (0, 1]
        u = (raw >> np.uint64(11)).astype(np.float64) * 2.0**-53
        assert u.min() >= 0 and u.max() < 1


@pytest.mark.parametrize("seed", [1, 42, 123456789])
def test_generators_match_the_oracle(seed):
    assert np.array_equal(W.uniform_i32(seed, -1000, 1000, 50001), O.gen_uniform_i32(seed, -1000, 1000, 50001))
    assert np.array_equal(W.uniform_i32(seed, -2**31, 2**31 - 1, 999), O.gen_uniform_i32(seed, -2**31, 2**31 - 1, 999))
    assert np.array_equal(W.uniform_f32(seed, -1.0, 1.0, 40000), O.gen_uniform_f32(seed, -1.0, 1.0, 40000))
    assert np.array_equal(W.normal_f32(seed, 40001), O.gen_normal_f32(seed, 40001))
    assert np.array_equal(W.zipf_i32(seed, 1.1, 4096, 30000), O.gen_zipf_i32(seed, 1.1, 4096, 30000))
    for nseg, maxlen, total in ((1 << 12, 16384, 1 << 22), (1000, 5, 1200), (7, 1000, 7)):
        assert np.array_equal(W.segment_lengths(seed, nseg, maxlen, total),
                              O.gen_segment_lengths(seed, nseg, maxlen, total))


def test_distributions():
    x = W.normal_f32(42, 1 << 20).astype(np.float64)
    assert abs(x.mean()) < 0.005 and abs(x.std() - 1) < 0.005
    # Box-Muller pairs: cos and sin of one angle, so x0^2 + x1^2 = -2 ln u1 > 0
    assert np.all(x[0::2] ** 2 + x[1::2] ** 2 > 0)
    f = W.uniform_f32(7, 2.0, 3.0, 1 << 18)
    assert f.min() >= 2.0 and f.max() < 3.0 and abs(f.mean() - 2.5) < 0.005
    u = W.uniform_i32(42, -1000, 1000, 1 << 20)
    assert abs(u.mean()) < 2 and u.min() == -1000 and u.max() == 1000
    # C2's lengths: log-uniform before scaling, so log L is roughly uniform
    lens = W.segment_lengths(42, 1 << 16, 16384, 1 << 26)
    assert lens.sum() == 1 << 26 and lens.min() >= 1
    z = W.zipf_i32(3, 1.1, 4096, 1 << 20)
    p = np.arange(1, 4097, dtype=np.float64) ** -1.1
    p /= p.sum()
    counts = np.bincount(z, minlength=4096)
    assert abs(counts[0] / z.size - p[0]) < 0.003 and abs(counts[1] / z.size - p[1]) < 0.003
