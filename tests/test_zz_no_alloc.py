"""amdk_overlap_noalloc (split-K partials in the library's fixed per-GPU
array) against amdk_overlap (context split-K buffer): two calls in a row,
the second smaller, in both precision modes; and the size limit."""
import numpy as np
import pytest

from amdkiit_amd.context import KSContext
from amdkiit_amd.lib import AmdkError, ERR_ARG
from tests.common import random_bands, small_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_overlap_noalloc_matches_overlap(prec):
    pw = small_setup(10.0, 12.0)
    ctx = KSContext(pw, 40, prec)
    for n1, n2, seed in ((40, 33, 1), (9, 5, 2)):
        print(f"case=overlap_noalloc prec={prec} n1={n1} n2={n2} npw={pw.ngpw}", flush=True)
        ca = random_bands(pw, n1, seed=seed)
        cb = random_bands(pw, n2, seed=seed + 10)
        ref = ctx.overlap(ca, cb)
        got = ctx.overlap_noalloc(ca, cb)
        assert got.shape == (n1, n2)
        # the same kernels and split count: identical sums
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        direct = 2.0 * (ca.conj() @ cb.T).real - np.outer(ca[:, 0].real, cb[:, 0].real)
        assert np.abs(got - direct).max() < 1e-12 * np.abs(direct).max()
    ctx.close()


def test_overlap_noalloc_rejects_products_beyond_its_scratch():
    pw = small_setup(6.0, 4.0)
    ctx = KSContext(pw, 2100, "fp64")
    c = random_bands(pw, 2049, seed=3)
    print(f"case=overlap_noalloc_limit npw={pw.ngpw} n=2049", flush=True)
    with pytest.raises(AmdkError) as ei:
        ctx.overlap_noalloc(c, c)
    assert ei.value.code == ERR_ARG
    assert "allocation-free limit" in str(ei.value)
    ctx.close()
