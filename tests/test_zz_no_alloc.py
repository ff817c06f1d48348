"""dq_marginal_probs_noalloc (no device allocation; work space in a fixed
per-GPU array) against dq_marginal_probs, through the C ABI.

Both run the same kernels with the same reduction shapes, so the values are
bit-identical.  Two calls in a row, the second with a smaller input, show
that the shared array keeps no state that changes a later result.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

capi = pytest.importorskip("dqsim_bosonic_amd.capi")


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


@pytest.mark.parametrize("dtype", [capi.C128, capi.C64])
def test_noalloc_marginal_equals_marginal(dtype):
    cases = [
        (20, [19, 3, 7, 0, 11, 2, 15, 8, 1, 18, 5, 13, 9, 4, 17, 6, 12, 10]),  # 18 qubits, tree path: 4 MiB + 512 B
        (20, [7, 18, 0, 12, 3, 15, 9, 1, 16, 5, 11, 2, 14, 6, 17, 10, 4, 13, 8]),  # 19 qubits, 2 terms per bin: 4 MiB + 512 B
        (12, [3, 9, 1]),                                                      # smaller: 2^9 terms per bin
        (10, list(range(9, -1, -1))),                                        # k = n: elementwise path
        (14, [2, 13, 5, 7, 0, 11, 9, 12, 1]),                                # 2^5 terms per bin: sequential path
        (8, []),                                                             # k = 0: the norm
    ]
    for i, (n, qs) in enumerate(cases):
        print(f"[noalloc] dtype={dtype} n={n} k={len(qs)}", flush=True)
        st = capi.State(n, dtype)
        st.set(random_state(n, 10 + i))
        st.synchronize()
        want = st.marginal_probs(qs)
        got = st.marginal_probs_noalloc(qs)
        assert got.shape == want.shape
        assert np.array_equal(got, want), (n, qs)
        # a second call in a row, with a smaller input
        sub = qs[: max(0, len(qs) - 3)]
        assert np.array_equal(st.marginal_probs_noalloc(sub), st.marginal_probs(sub))
        st.close()


def test_noalloc_marginal_rejects_what_it_cannot_do():
    st = capi.State(20)
    st.set(random_state(20, 3))
    st.synchronize()
    with pytest.raises(capi.DqError) as e:
        st.marginal_probs_noalloc(list(range(20)))  # 2^20 bins: 8 MiB, beyond the work space
    assert e.value.code == capi.ERR_ARG and "maximum" in str(e.value)
    st.apply_1q(np.array([[0, 1], [1, 0]], dtype=np.complex128), 4)  # queued, not yet applied
    with pytest.raises(capi.DqError) as e:
        st.marginal_probs_noalloc([4])
    assert e.value.code == capi.ERR_ARG and "queued" in str(e.value)
    st.synchronize()
    got = st.marginal_probs_noalloc([4])
    assert np.array_equal(got, st.marginal_probs([4]))
    st.close()
