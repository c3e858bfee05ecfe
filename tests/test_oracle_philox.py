"""The numpy Philox4x32-10 (oracle/philox.py) against the known-answer
vectors of the Random123 distribution (kat_vectors: philox4x32 with 10
rounds), and the start-field layout: every word lands on its own component,
values in [-1/2, 1/2)."""
import numpy as np

import philox as ph


def _hex(c):
    return [int(v) for v in c]


def test_known_answers():
    assert _hex(ph.philox4x32_10(0, 0, 0, 0, 0, 0)) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    f = 0xffffffff
    assert _hex(ph.philox4x32_10(f, f, f, f, f, f)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert _hex(ph.philox4x32_10(0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0)) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_start_field_layout():
    key = (0.25, 0.75)
    s = ph.start_field(8, key)
    k0, k1 = 1 << 30, 3 << 30
    for x in (0, 7):
        for j in range(6):
            w = ph.philox4x32_10(x, j, 0, 0, k0, k1)
            for i in range(4):
                assert s[x, 4 * j + i] == np.float32(int(w[i]) >> 8) * np.float32(2.0 ** -24) - np.float32(0.5)
    assert (s >= -0.5).all() and (s < 0.5).all()
