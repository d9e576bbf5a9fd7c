"""The allocation-free entry point hykkt_ruiz_scale_noalloc through the C ABI
against hykkt_ruiz_scale (RuizClass.cpp:54-98): the same bits on two calls in
a row, the second with a smaller system (the scratch keeps the first call's
contents), for two sweeps and for none; and a clear error for a system the
scratch cannot hold (checked before any device work).  Needs an MI355X."""
import numpy as np
import pytest
import scipy.sparse as sp

import hykkt_amd as H

pytestmark = pytest.mark.gpu


def ruiz_inputs(n, m2, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=4.0 / n, random_state=np.random.RandomState(seed))
    A = (A + A.T + sp.diags(rng.uniform(1, 10, n))).tocsr()
    A.sort_indices()
    J = sp.random(m2, n, density=6.0 / n, random_state=np.random.RandomState(seed + 1)).tocsr()
    J = (J + sp.csr_matrix((np.ones(m2), (np.arange(m2), np.arange(m2) % n)), shape=(m2, n)))
    J = J.tocsr()
    J.sort_indices()
    Jt = J.T.tocsr()
    Jt.sort_indices()
    trip = lambda M: (M.indptr.astype(np.int32), M.indices.astype(np.int32),
                      M.data.astype(np.float64) * rng.choice([-3.0, 0.5, 7.0], size=M.nnz))
    Hh, Jh = trip(A), trip(J)
    Jth = (Jt.indptr.astype(np.int32), Jt.indices.astype(np.int32), None)
    # Jt holds the same values as J, in transposed order
    Jtv = sp.csr_matrix((Jh[2], Jh[1], Jh[0]), shape=(m2, n)).T.tocsr()
    Jtv.sort_indices()
    Jth = (Jtv.indptr.astype(np.int32), Jtv.indices.astype(np.int32), Jtv.data)
    return Hh, Jth, Jh, rng.standard_normal(n), rng.standard_normal(m2)


@pytest.mark.parametrize("its", [2, 0])
def test_noalloc_same_bits_twice(its):
    for n, m2, seed in ((3000, 1400, 1), (700, 300, 2)):
        args = ruiz_inputs(n, m2, seed)
        want = H.ruiz_scale(its, *args)
        got = H.ruiz_scale(its, *args, noalloc=True)
        for a, b in zip(want, got):
            assert np.array_equal(a, b), (n, its)


def test_noalloc_refuses_a_system_beyond_the_scratch():
    n = 1_500_000
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.ones(n)
    with pytest.raises(H.HykktError) as ex:
        H.ruiz_scale(2, (rp, ci, v), (rp, np.zeros(0, np.int32), np.zeros(0)),
                     (np.zeros(2, np.int32), np.zeros(0, np.int32), np.zeros(0)),
                     np.ones(n), np.ones(1), noalloc=True)
    assert ex.value.code == -1 and "device scratch" in str(ex.value)
