"""Pin of random_shfl, the reference CUDA harness's row shuffle
(cuda/force_cuda.cu:255-263: one std::mt19937(10) stream, std::shuffle of
every CSR row in row order).

The permutation is restated here from published algorithms, independently of
the C++ the product and the oracle call:
  * the MT19937 stream: numpy's MT19937 with its legacy integer seeding is
    Matsumoto-Nishimura's init_genrand, the seeding std::mt19937(seed) uses;
    checked against the C++ standard's required value (the 10000th output of
    a default-seeded std::mt19937 is 4123659995, [rand.predef]);
  * std::uniform_int_distribution of libstdc++ 11 (this image's, which the
    product and oracle link): for a 32-bit generator, Lemire's nearly
    divisionless method (Lemire 2019, ACM TOMACS 29(1), "Fast random integer
    generation in an interval");
  * std::shuffle of libstdc++ 11: Fisher-Yates from the front, position i
    swapped with a uniform draw in [0, i]; when the generator's range covers
    the square of the row length, two positions come from one draw x in
    [0, (i+1)(i+2)) as x / (i+2) and x % (i+2), with a single [0, 1] draw
    first when the row length is even.
An older libstdc++ (division-based uniform_int_distribution, one draw per
position) gives another permutation of each row; the shuffle only reorders
each row's sum, which moves forces by about 1e-16 relative. CPU only.
"""
import numpy as np
import pytest

import lj_gpu_amd as G
import oracle as O
from helpers import periodic, reference_mode

MASK32 = (1 << 32) - 1


class MTStream:
    """32-bit outputs of MT19937 seeded as std::mt19937(seed)."""

    def __init__(self, seed: int):
        self.bg = np.random.MT19937(0)
        self.bg._legacy_seeding(seed)
        self.buf = np.zeros(0, np.uint64)
        self.pos = 0

    def __call__(self) -> int:
        if self.pos == self.buf.size:
            self.buf = self.bg.random_raw(4096)
            self.pos = 0
        v = int(self.buf[self.pos])
        self.pos += 1
        return v


def uniform_int(draw, a: int, b: int) -> int:
    """libstdc++ 11 uniform_int_distribution{a, b} over a 32-bit generator
    (b - a < 2^32 - 1): Lemire's method with 64-bit products."""
    erange = (b - a + 1) & MASK32
    product = draw() * erange
    low = product & MASK32
    if low < erange:
        threshold = ((1 << 32) - erange) % erange
        while low < threshold:
            product = draw() * erange
            low = product & MASK32
    return a + (product >> 32)


def shuffle(row: list, draw) -> None:
    n = len(row)
    if n == 0:
        return
    if MASK32 // n >= n:
        i = 1
        if n % 2 == 0:
            j = uniform_int(draw, 0, 1)
            row[i], row[j] = row[j], row[i]
            i += 1
        while i != n:
            r = i + 1
            x = uniform_int(draw, 0, r * (r + 1) - 1)
            a, b = x // (r + 1), x % (r + 1)
            row[i], row[a] = row[a], row[i]
            i += 1
            row[i], row[b] = row[b], row[i]
            i += 1
    else:
        for i in range(1, n):
            j = uniform_int(draw, 0, i)
            row[i], row[j] = row[j], row[i]


def restated_random_shfl(csr, seed: int = 10, reseed_per_row: bool = False) -> np.ndarray:
    out = csr.sorted_list.copy()
    draw = MTStream(seed)
    for i in range(csr.n):
        if reseed_per_row:
            draw = MTStream(seed)
        a = int(csr.pointer[i])
        row = out[a: a + int(csr.number_of_partners[i])].tolist()
        shuffle(row, draw)
        out[a: a + len(row)] = row
    return out


def test_mt19937_stream_is_the_standard_one():
    draw = MTStream(5489)
    for _ in range(9999):
        draw()
    assert draw() == 4123659995
    # The legacy seeding is init_genrand: x0 = seed, then
    # x_k = 1812433253 * (x_{k-1} ^ (x_{k-1} >> 30)) + k (mod 2^32), k < 624.
    state = np.random.RandomState(10).get_state()[1]
    x = [10]
    for k in range(1, 624):
        x.append((1812433253 * (x[-1] ^ (x[-1] >> 30)) + k) & MASK32)
    np.testing.assert_array_equal(state, np.array(x, np.uint32))


def test_uniform_int_is_unbiased_and_in_range():
    draw = MTStream(1)
    for b in (1, 2, 6, 142 * 143 - 1, 1000003):
        v = [uniform_int(draw, 0, b) for _ in range(2000)]
        assert min(v) >= 0 and max(v) <= b
    # Lemire's rejection: a range that does not divide 2^32 rejects the
    # lowest (2^32 mod range) products; with range 3 the threshold is 1.
    assert ((1 << 32) - 3) % 3 == 1


@pytest.mark.parametrize("which", ["periodic_full", "reference_half"])
def test_random_shfl_matches_the_restated_permutation(which):
    if which == "periodic_full":
        q, box, _, csr = periodic(6, 1.0)
    else:
        q, csr = reference_mode(0.5)
    want = restated_random_shfl(csr, 10)
    got_product = G.random_shfl(csr, 10).sorted_list
    got_oracle = O.random_shfl(O.CSR(csr.number_of_partners, csr.pointer, csr.sorted_list,
                                     "full"), 10).sorted_list
    np.testing.assert_array_equal(got_product, want)
    np.testing.assert_array_equal(got_oracle, want)
    # The pin is sharp: a per-row reseed or another seed gives another list.
    assert not np.array_equal(restated_random_shfl(csr, 10, reseed_per_row=True), want)
    assert not np.array_equal(restated_random_shfl(csr, 11), want)


def test_restated_shuffle_small_rows_by_hand():
    """Rows of length 0 and 1 draw nothing; a row of 2 draws one [0, 1]
    value; a row of 3 draws one value in [0, 6) for positions 1 and 2."""
    for n, draws in ((0, 0), (1, 0), (2, 1), (3, 1), (4, 2), (5, 2)):
        count = [0]
        inner = MTStream(3)

        def draw():
            count[0] += 1
            return inner()
        row = list(range(n))
        shuffle(row, draw)
        assert sorted(row) == list(range(n))
        assert count[0] >= draws  # rejections may add draws, never remove them
