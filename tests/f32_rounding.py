"""FP64 -> FP32 conversion cases whose results the IEEE 754 rules fix
(round to nearest, ties to even; subnormals; overflow to infinity), written
as (double, expected float32 bit pattern).  Used to pin the conversion of
the oracle (oracle_convert_d2f) and of the product (convertD2F, the
reference's src_cpp/loadmatrix.cpp:252-342, which casts element by
element)."""
import numpy as np

FLT_MAX = (2.0 - 2.0 ** -23) * 2.0 ** 127
CASES = [
    (1.0, 0x3F800000),
    (1.0 + 2.0 ** -24, 0x3F800000),                # tie: 1 is even
    (1.0 + 3 * 2.0 ** -24, 0x3F800002),            # tie: rounds up to the even mantissa
    (1.0 + 2.0 ** -24 + 2.0 ** -52, 0x3F800001),   # just above the tie
    (1.0 + 2.0 ** -24 - 2.0 ** -52, 0x3F800000),   # just below the tie
    (-(1.0 + 3 * 2.0 ** -24), 0xBF800002),
    (0.1, 0x3DCCCCCD),
    (-2.5, 0xC0200000),
    (2.0 ** -126, 0x00800000),                     # smallest normal
    (2.0 ** -149, 0x00000001),                     # smallest subnormal
    (2.0 ** -150, 0x00000000),                     # tie between 0 and 2^-149: 0 is even
    (3 * 2.0 ** -150, 0x00000002),                 # tie between 1 and 2 units: 2 is even
    (2.0 ** -150 + 2.0 ** -160, 0x00000001),
    (1e-50, 0x00000000),
    (-1e-50, 0x80000000),                          # sign kept through underflow
    (FLT_MAX, 0x7F7FFFFF),
    (FLT_MAX + 2.0 ** 102, 0x7F7FFFFF),            # below the tie with 2^128
    (FLT_MAX + 2.0 ** 103, 0x7F800000),            # tie: the even neighbour is 2^128 -> inf
    (1e39, 0x7F800000),
    (-1e39, 0xFF800000),
    (float("inf"), 0x7F800000),
    (float("-inf"), 0xFF800000),
    (0.0, 0x00000000),
    (-0.0, 0x80000000),
]


def case_values():
    return np.array([d for d, _ in CASES], dtype=np.float64)


def case_bits():
    return np.array([b for _, b in CASES], dtype=np.uint32)


def random_values(n, seed):
    """Doubles over the whole float32 range and beyond, both signs."""
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, n) * np.exp2(rng.integers(-160, 140, n))).astype(np.float64)


def check(src, got):
    """Every element of got is src rounded to the nearest float32 (NaN -> NaN)."""
    src = np.asarray(src, dtype=np.float64)
    got = np.asarray(got, dtype=np.float32)
    nan = np.isnan(src)
    assert np.array_equal(np.isnan(got), nan)
    with np.errstate(over="ignore"):
        want = src[~nan].astype(np.float32)
    assert np.array_equal(got[~nan].view(np.uint32), want.view(np.uint32))
