"""Pencil decomposition and transposes (src/decomp_2d.cc:41-589).

CPU: the oracle's decomposition is pinned by hand-worked distributions (the
2DECOMP rule that the last blocks take the remainder) and by the invariant that
defines a transpose -- every value lands at the same global (x, y, z) in the
destination pencil -- checked against an independent numpy scatter/gather.
The library's host bookkeeping must agree with the oracle rank by rank.

GPU: the four transpose entry points, through the C-ABI on device buffers,
must reproduce the oracle's mem_split -> Alltoallv -> mem_merge bit for bit.
"""
import numpy as np
import pytest

import powerllel_channel_sycl_amd as pll

# (nx, ny, nz, p_row, p_col): even, uneven, p_row == p_col, single-rank and
# degenerate 1 x P / P x 1 grids.
GRIDS = [
    (8, 8, 8, 2, 2),
    (10, 7, 9, 3, 2),
    (12, 9, 11, 3, 3),
    (17, 13, 5, 4, 5),
    (6, 5, 4, 1, 1),
    (16, 12, 10, 1, 4),
    (16, 12, 10, 4, 1),
    (48, 40, 36, 2, 8),
]
KINDS = ["x_to_y", "y_to_x", "y_to_z", "z_to_y"]
SRC = {"x_to_y": 0, "y_to_x": 1, "y_to_z": 1, "z_to_y": 2}
DST = {"x_to_y": 1, "y_to_x": 0, "y_to_z": 2, "z_to_y": 1}


def _pack(dec, pencil, glob):
    """Every rank's pencil of a global (nz, ny, nx) array, concatenated in rank order."""
    parts = []
    for r in range(dec.nranks):
        st, sz, _ = dec.pencil(pencil, r)
        i0, j0, k0 = (s - 1 for s in st)
        parts.append(glob[k0:k0 + sz[2], j0:j0 + sz[1], i0:i0 + sz[0]].ravel())
    return np.concatenate(parts)


def _unpack(dec, pencil, flat, shape):
    out = np.full(shape, np.nan)
    for r in range(dec.nranks):
        st, sz, off = dec.pencil(pencil, r)
        i0, j0, k0 = (s - 1 for s in st)
        n = sz[0] * sz[1] * sz[2]
        out[k0:k0 + sz[2], j0:j0 + sz[1], i0:i0 + sz[0]] = \
            flat[off:off + n].reshape(sz[2], sz[1], sz[0])
    return out


# ----------------------------------------------------------------------------- CPU


def test_distribute_golden(oracle_lib):
    # distribute(10, 3): sizes 3, 3, 4 starting at 1, 4, 7; distribute(7, 4): 1, 2, 2, 2.
    d = oracle_lib.OracleDecomp(10, 7, 4, 3, 4)
    assert [d.pencil(0, r * 4)[0][1] for r in range(3)] == [1, 3, 5]      # y over p_row (7/3)
    assert [d.pencil(0, r * 4)[1][1] for r in range(3)] == [2, 2, 3]
    assert [d.pencil(1, r * 4)[0][0] for r in range(3)] == [1, 4, 7]      # x over p_row (10/3)
    assert [d.pencil(1, r * 4)[1][0] for r in range(3)] == [3, 3, 4]
    assert [d.pencil(2, c)[1][1] for c in range(4)] == [1, 2, 2, 2]       # y over p_col (7/4)
    assert [d.pencil(0, c)[1][2] for c in range(4)] == [1, 1, 1, 1]       # z over p_col (4/4)


@pytest.mark.parametrize("grid", GRIDS)
def test_oracle_transpose_is_the_global_permutation(oracle_lib, grid):
    nx, ny, nz, pr, pc = grid
    d = oracle_lib.OracleDecomp(*grid)
    rng = np.random.default_rng(sum(grid))
    glob = rng.standard_normal((nz, ny, nx))
    for pencil in range(3):
        assert d.total(pencil) == nx * ny * nz
        back = _unpack(d, pencil, _pack(d, pencil, glob), glob.shape)
        assert np.array_equal(back, glob)
    for kind in KINDS:
        out = d.transpose(kind, _pack(d, SRC[kind], glob))
        assert np.array_equal(out, _pack(d, DST[kind], glob)), kind


@pytest.mark.parametrize("grid", GRIDS)
def test_library_bookkeeping_matches_oracle(oracle_lib, grid):
    d = pll.Decomp(*grid)
    o = oracle_lib.OracleDecomp(*grid)
    for pencil in range(3):
        assert d.total(pencil) == o.total(pencil)
        for r in range(d.nranks):
            assert d.pencil(pencil, r) == o.pencil(pencil, r), (pencil, r)
    d.close()


def test_block_split_sweep(oracle_lib):
    # Every split of n over p ranks (the x extent over p_row in a y-pencil): counts are
    # floor(n/p) with the last n % p ranks one longer, starts are contiguous from 1,
    # and library and oracle agree rank by rank.
    for n in range(1, 41):
        for p in range(1, min(n, 9) + 1):
            d = pll.Decomp(n, n, n, p, 1)
            o = oracle_lib.OracleDecomp(n, n, n, p, 1)
            starts = [d.pencil(1, r)[0][0] for r in range(p)]
            counts = [d.pencil(1, r)[1][0] for r in range(p)]
            want = [n // p + (1 if r >= p - n % p else 0) for r in range(p)]
            assert counts == want, (n, p)
            assert starts == [1 + sum(want[:r]) for r in range(p)], (n, p)
            assert all(d.pencil(1, r) == o.pencil(1, r) for r in range(p)), (n, p)
            d.close()


def test_invalid_processor_grid():
    # src/decomp_2d.cc:117-124: min(nx,ny) >= p_row and min(ny,nz) >= p_col
    with pytest.raises(pll.PllError, match="Invalid 2D processor grid"):
        pll.Decomp(4, 4, 2, 2, 3)
    with pytest.raises(pll.PllError):
        pll.Decomp(4, 4, 4, 0, 1)
    d = pll.Decomp(4, 4, 4, 2, 2)
    with pytest.raises(pll.PllError):
        d.pencil(3, 0)
    with pytest.raises(pll.PllError):
        d.pencil(0, 4)


# ----------------------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("grid", GRIDS + [(768, 384, 24, 4, 6)])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_transposes_bitexact(oracle_lib, grid, dtype):
    import torch
    print(f"[decomp] grid={grid} dtype={dtype}", flush=True)
    nx, ny, nz, pr, pc = grid
    d = pll.Decomp(*grid)
    o = oracle_lib.OracleDecomp(*grid)
    rng = np.random.default_rng(7 + sum(grid))
    glob = rng.standard_normal((nz, ny, nx))
    if dtype == "f32":
        glob = glob.astype(np.float32).astype(np.float64)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    esz = 8 if dtype == "f64" else 4
    for kind in KINDS:
        src = _pack(o, SRC[kind], glob)
        want = o.transpose(kind, src)
        s = torch.from_numpy(src).to(tdt).to("cuda:0")
        t = torch.full((d.total(DST[kind]),), float("nan"), dtype=tdt, device="cuda:0")
        d.transpose(kind, s.data_ptr(), t.data_ptr(), esz)
        got = t.cpu().double().numpy()
        assert np.array_equal(got, want), (kind, grid)
    # round trip x -> y -> z -> y -> x
    x = torch.from_numpy(_pack(o, 0, glob)).to(tdt).to("cuda:0")
    y = torch.empty(d.total(1), dtype=tdt, device="cuda:0")
    z = torch.empty(d.total(2), dtype=tdt, device="cuda:0")
    d.transpose("x_to_y", x.data_ptr(), y.data_ptr(), esz)
    d.transpose("y_to_z", y.data_ptr(), z.data_ptr(), esz)
    y.fill_(float("nan"))
    d.transpose("z_to_y", z.data_ptr(), y.data_ptr(), esz)
    x2 = torch.empty_like(x)
    d.transpose("y_to_x", y.data_ptr(), x2.data_ptr(), esz)
    assert torch.equal(x, x2)
    d.close()
