"""The m-major compact rotation layout: a bijection onto [0, nRc), and every
row the drift couples to (l, m, n) -- (l, m, n+-1), (l+-1, m, n+-1) -- within
2B-1 rows of it (the property the fused kernel's cache locality relies on)."""
import pytest

import hpup


@pytest.mark.parametrize("B", [1, 2, 5, 15, 30])
def test_ridx_bijection_and_locality(B):
    nrc = B * (2 * B - 1) * (2 * B + 1) // 3
    seen = set()
    worst = 0
    for l in range(B):
        for m in range(-l, l + 1):
            for n in range(-l, l + 1):
                r = hpup.ridx(B, l, m, n)
                seen.add(r)
                for dl in (-1, 0, 1):
                    for dn in (-1, 0, 1):
                        l1, n1 = l + dl, n + dn
                        if 0 <= l1 < B and abs(m) <= l1 and abs(n1) <= l1:
                            worst = max(worst, abs(hpup.ridx(B, l1, m, n1) - r))
    assert seen == set(range(nrc))
    assert worst <= max(2 * B - 1, 0)
