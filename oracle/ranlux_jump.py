"""Skip-ahead of the ranlxd streams -- test infrastructure (the oracle), not
part of the library.

ranlxd (random/ranlxd.c, random/ranlux_common.c) runs four independent
lanes of the subtract-with-borrow recursion in base b = 2^48,

    x_n = x_{n-5} - x_{n-12} - c_{n-1}  (mod b),  c_n = 1 if the difference
    was negative,

advanced by pr = 202 (level 1) or 397 (level 2) steps per refill, after which
the last 12 words of the four lanes are handed out as 48 doubles x 2^-48 in
the order word (oldest first) major, lane minor (rlx_update and
rlx_convertd, ranlux_common.c:569-608, 784-800; ranlxd, ranlxd.c:129-162).
Each lane is the linear congruential generator

    Z_n = a Z_{n-1} mod m,  m = b^12 - b^5 + 1,  a = b^-1 mod m

with Z_n = sum_k x_{n-11+k} b^k - sum_{k<5} x_{n-4+k} b^k + c_n, and the words
are the base-b digits x_n = floor(b Z_n / m), Z_{n-1} = b Z_n mod m (A.
Sibidanov, Comput. Phys. Commun. 221 (2017) 299).  A jump by N steps is a
product with a^N mod m.  Python integers hold the residues exactly.
"""

B = 1 << 48
M = B ** 12 - B ** 5 + 1
A = pow(B, -1, M)


def seed_lanes(seed):
    """The seeded words of the four lanes (rlx_init with flag 1,
    ranlux_common.c:117-171): a 31-bit shift register gives 24-bit numbers,
    24 per lane; number k of lane i is complemented unless k % 4 == i; the
    even numbers are the low and the odd ones the high halves of the 12
    48-bit words.  Returns words[lane][12] in chronological order (the
    recursion starts at word 0) and the four borrows (zero)."""
    bits = [(seed >> k) & 1 for k in range(31)]
    p, q = 0, 18
    words = [[0] * 12 for _ in range(4)]
    for lane in range(4):
        for k in range(24):
            w = 0
            for _ in range(24):
                w = 2 * w + bits[p]
                bits[p] = (bits[p] + bits[q]) & 1
                p = (p + 1) % 31
                q = (q + 1) % 31
            if k % 4 != lane:
                w = 0xffffff - w
            if k & 1:
                words[lane][k // 2] += w << 24
            else:
                words[lane][k // 2] = w
    return words, [0, 0, 0, 0]


def step(words, carry, n):
    """n steps of the recursion on one lane (chronological words)."""
    w = list(words)
    c = carry
    for _ in range(n):
        d = w[-5] - w[0] - c
        c = 1 if d < 0 else 0
        w = w[1:] + [d % B]
    return w, c


def to_z(words, carry):
    """Residue of a lane state."""
    p = sum(x << (48 * k) for k, x in enumerate(words))
    q = sum(x << (48 * k) for k, x in enumerate(words[7:]))
    return (p - q + carry) % M


def from_z(z):
    """Lane state (chronological words, borrow) of a residue."""
    xs = []
    y = z
    for _ in range(12):
        xs.append((B * y) // M)
        y = (B * y) % M
    w = xs[::-1]
    c = z - (sum(x << (48 * k) for k, x in enumerate(w)) - sum(x << (48 * k) for k, x in enumerate(w[7:])))
    c %= M
    assert c in (0, 1)
    return w, c


def jump(words, carry, n):
    """The lane state n steps on."""
    return from_z(pow(A, n, M) * to_z(words, carry) % M)


def uniforms(seed, level, n0, n1):
    """Numbers n0 .. n1-1 of a freshly seeded ranlxd stream (rlxd_init with
    this level, ranlxd.c:98-125), each computed by a jump to its refill."""
    pr = 202 if level == 1 else 397
    lanes, carries = seed_lanes(seed)
    out = []
    cache = {}
    for n in range(n0, n1):
        j = n // 48 + 1          # refill that hands out number n
        m, l = (n % 48) // 4, n % 4
        key = (j, l)
        if key not in cache:
            cache[key] = jump(lanes[l], carries[l], j * pr)[0]
        out.append(cache[key][m] * 2.0 ** -48)
    return out
