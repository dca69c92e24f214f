"""Pins the oracle's sampling layer (a1-a4) against the reference's
known-answer tests (src/pathtrace.rs:2866-2908)."""
import numpy as np


def to_unit(v):
    return np.float32((v >> 8) * (1.0 / 16777216.0))


def test_sobol_unscrambled_dims_match_published_points(oracle):
    # Joe & Kuo's first eight points in Gray-code order (src/pathtrace.rs:2871-2880).
    table = [
        [0.0, 0.0, 0.0, 0.0],
        [0.5, 0.5, 0.5, 0.5],
        [0.75, 0.25, 0.25, 0.25],
        [0.25, 0.75, 0.75, 0.75],
        [0.375, 0.375, 0.625, 0.875],
        [0.875, 0.875, 0.125, 0.375],
        [0.625, 0.125, 0.875, 0.625],
        [0.125, 0.625, 0.375, 0.125],
    ]
    for i, row in enumerate(table):
        g = i ^ (i >> 1)
        got = [float(to_unit(oracle.sobol(g, d))) for d in range(4)]
        assert got == row, (i, got)
    # Dimensions 0 and 1 form a (0,2)-sequence.
    cells = np.zeros(16, int)
    for i in range(16):
        x = int(float(to_unit(oracle.sobol(i, 0))) * 4)
        y = int(float(to_unit(oracle.sobol(i, 1))) * 4)
        cells[y * 4 + x] += 1
    assert (cells == 1).all()


def test_direction_numbers_are_joe_kuo():
    # m_k of dims 1..3 from (s, a, m) = (1,0,[1]), (2,1,[1,3]), (3,1,[1,3,1]):
    # the published first direction numbers, scaled by 2^(31-k).
    import oracle as O
    m = O.sobol_matrices()
    assert [int(x) for x in m[0][:4]] == [1 << 31, 1 << 30, 1 << 29, 1 << 28]
    # dim 1 (x+1): m = 1,1,1,1,... -> v_k = 2^(31-k) * m_k with recurrence m_k = m_{k-1} ^ (2 m_{k-1}).
    exp_m1 = [1]
    for k in range(1, 8):
        exp_m1.append(exp_m1[-1] ^ (exp_m1[-1] << 1))
    assert [int(m[1][k]) >> (31 - k) for k in range(8)] == exp_m1
    # dim 2 (x^2+x+1): m = 1, 3, then m_k = 2 m_{k-1} ^ 4 m_{k-2} ^ m_{k-2}.
    exp_m2 = [1, 3]
    for k in range(2, 8):
        exp_m2.append((exp_m2[k - 1] << 1) ^ (exp_m2[k - 2] << 2) ^ exp_m2[k - 2])
    assert [int(m[2][k]) >> (31 - k) for k in range(8)] == exp_m2
    # dim 3 (x^3+x+1): m = 1, 3, 1, then m_k = 2 m_{k-2} ^ 8 m_{k-3} ^ m_{k-3}.
    exp_m3 = [1, 3, 1]
    for k in range(3, 8):
        exp_m3.append((exp_m3[k - 2] << 2) ^ (exp_m3[k - 3] << 3) ^ exp_m3[k - 3])
    assert [int(m[3][k]) >> (31 - k) for k in range(8)] == exp_m3


def test_owen_scrambling_preserves_stratification(oracle):
    for seed in [1, 77, 0xDEADBEEF]:
        p = oracle.sobol_owen_4d(np.arange(64, dtype=np.uint32), np.full(64, seed, np.uint32))
        assert ((p >= 0) & (p < 1)).all()
        cells = np.zeros(64, int)
        for q in p:
            cells[int(q[0] * 8) * 8 + int(q[1] * 8)] += 1
        assert (cells == 1).all(), seed
        # Every 1-D projection is stratified too (one point per 1/64 interval).
        for d in range(4):
            assert sorted((p[:, d] * 64).astype(int)) == list(range(64))


def test_scrambled_points_are_24_bit_dyadic(oracle):
    p = oracle.sobol_owen_4d(np.arange(1000, dtype=np.uint32), np.full(1000, 12345, np.uint32))
    q = p.astype(np.float64) * (1 << 24)
    assert np.array_equal(q, np.floor(q))
