"""wf_quotients_static, the entry point that allocates nothing (its staging area is a static
device array), against wf_quotients_base through the C ABI: two calls in a row, the second
smaller, so that stale values of the first call in the shared array would show; sizes above
the array are refused with WF_E_LIMIT."""
import random

import pytest

from mathpub_amd import wf

pytestmark = pytest.mark.gpu


def odd_values(n, seed):
    rng = random.Random(seed)
    out = [3, 1093, 3511, (1 << 63) + 29, (1 << 64) - 59, (1 << 64) - 1]
    while len(out) < n:
        bits = rng.randint(2, 64)
        x = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if x >= 3:
            out.append(x)
    return out[:n]


@pytest.mark.parametrize("arith,base", [(0, 2), (0, 3), (0, 4294967291), (1, 2)])
def test_static_equals_allocating_entry_point(arith, base):
    with wf.Engine(arith=arith) as e:
        for n, seed in ((wf.STATIC_QUOT_MAX, 1), (1000, 2)):
            ps = odd_values(n, seed)
            want = e.quotients(ps, base=base)
            got = e.quotients_static(ps, base=base)
            assert got == want
        for p, (q, r) in zip(ps[:50], got[:50]):
            assert r == pow(base, p - 1, p * p)


def test_static_limit_is_refused():
    with wf.Engine() as e:
        with pytest.raises(wf.WFError) as ex:
            e.quotients_static(odd_values(wf.STATIC_QUOT_MAX + 1, 3))
        assert ex.value.rc == -7
        assert e.quotients_static([]) == []
        assert e.quotients_static([1093]) == e.quotients([1093])
