"""Host side of the GPU Wigner table (gsh_wigner_size / gsh_wigner_offset and
gshtrans_amd.Wigner.index): sizes and per-block indices against the oracle's
GSHIndices (Indexing.h:72-150), block offsets against the reference's
construction -- sizes accumulated over Wigner::Indices() in order
(Wigner.h:116-132, 240-250) -- and argument checks.  No GPU needed."""
import itertools

import pytest

import gshtrans_amd as G
import oracle

CASES = [(20, 20, "All", 2), (20, 7, "All", 2), (20, 1, "All", 2), (20, 20, "NonNegative", 2),
         (20, 5, "NonNegative", 3), (9, 9, "Single", 4), (33, 40, "NonNegative", 0)]


def upper(nrange, nmax):
    return {"All": range(-nmax, nmax + 1), "NonNegative": range(0, nmax + 1),
            "Single": range(nmax, nmax + 1)}[nrange]


@pytest.mark.parametrize("mrange", ["All", "NonNegative"])
@pytest.mark.parametrize("lmax,mmax,nrange,nmax", CASES)
def test_sizes_and_offsets(lmax, mmax, nrange, nmax, mrange):
    L = G.library()
    ns = list(upper(nrange, nmax))
    ntheta = 5
    mr = G._MRANGE[mrange]
    size = L.gsh_wigner_size(lmax, mmax, ns[0], ns[-1], ntheta, mr)
    assert size == ntheta * sum(oracle.gsh_size(lmax, mmax, n, mrange) for n in ns)
    for storage, code in G.WIGNER_STORAGE.items():
        # Wigner::Indices(): ColumnMajor iterates (n, iTheta), RowMajor (iTheta, n).
        order = (itertools.product(ns, range(ntheta)) if storage == "ColumnMajor" else
                 ((n, i) for i, n in itertools.product(range(ntheta), ns)))
        acc = 0
        for n, i in order:
            assert L.gsh_wigner_offset(lmax, mmax, ns[0], ns[-1], ntheta, mr, code, n, i) == acc
            acc += oracle.gsh_size(lmax, mmax, n, mrange)
        assert acc == size


@pytest.mark.parametrize("mrange", ["All", "NonNegative"])
@pytest.mark.parametrize("lmax,mmax,nrange,nmax", CASES[:5])
def test_block_index_matches_oracle(lmax, mmax, nrange, nmax, mrange):
    w = G.Wigner.__new__(G.Wigner)
    w.lmax, w.mmax, w.mrange = lmax, mmax, mrange
    mm = min(lmax, mmax)
    for n in upper(nrange, nmax):
        for l in range(abs(n), lmax + 1):
            ms = range(-min(l, mm), min(l, mm) + 1) if mrange == "All" else range(0, min(l, mm) + 1)
            for m in ms:
                assert w.index(n, l, m) == oracle.gsh_index(lmax, mmax, n, l, m, mrange)
        assert w.block_size(n) == oracle.gsh_size(lmax, mmax, n, mrange)


def test_invalid_arguments():
    L = G.library()
    assert L.gsh_wigner_size(-1, 4, 0, 0, 3, 0) == -1         # lmax < 0
    assert L.gsh_wigner_size(4, -1, 0, 0, 3, 0) == -1         # mmax < 0
    assert L.gsh_wigner_size(4, 4, 0, 5, 3, 0) == -1          # |n| > lmax
    assert L.gsh_wigner_size(4, 4, 1, 0, 3, 0) == -1          # nmin > nmax
    assert L.gsh_wigner_size(4, 4, 0, 0, 0, 0) == -1          # no angles
    assert L.gsh_wigner_size(4, 4, 0, 0, 3, 7) == -1          # bad mrange
    assert L.gsh_wigner_offset(4, 4, 0, 1, 3, 0, 2, 0, 0) == -1  # bad storage
    assert L.gsh_wigner_offset(4, 4, 0, 1, 3, 0, 0, 2, 0) == -1  # n outside
    assert L.gsh_wigner_offset(4, 4, 0, 1, 3, 0, 0, 0, 3) == -1  # theta index outside
