"""Host restatement of the device random streams, for tests: Philox4x32-10
(Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11: 10 rounds of two 32x32->64 multiplies with M0 = 0xD2511F53 and
M1 = 0xCD9E8D57, the key bumped by the Weyl constants 0x9E3779B9 and
0xBB67AE85), the counter layout of dml_philox, the 53-bit uniform and the
Box-Muller pair of gauss2 (din-mol-li_amd/csrc/dml_internal.h).  Written
from the paper, not from the device source; pinned by the SC11 answer
vectors in test_philox_kat.py."""
import math

MASK = 0xFFFFFFFF
M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
RS_BROWN, RS_ERMAK, RS_DEPO, RS_CONTACT, RS_GCMC = 1, 2, 3, 4, 5


def philox4x32_10(ctr, key):
    x0, x1, x2, x3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * x0
        p1 = M1 * x2
        x0, x1, x2, x3 = ((p1 >> 32) ^ x1 ^ k0) & MASK, p1 & MASK, ((p0 >> 32) ^ x3 ^ k1) & MASK, p0 & MASK
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return x0, x1, x2, x3


def draw(seed, pid, step, site, k):
    """the four words of (replica seed, particle id, step, site, draw k)"""
    return philox4x32_10((pid & MASK, step & MASK, ((site << 24) | (k & 0xFFFFFF)) & MASK, 0x6D6C6921),
                         (seed & MASK, (seed >> 32) & MASK))


def u01(a, b):
    """53 bits of two words, centred in their ulp: a uniform in (0, 1)"""
    return ((((a << 32) | b) >> 11) + 0.5) * (1.0 / 9007199254740992.0)


def gauss2(w):
    """Box-Muller pair of one Philox draw"""
    u1 = u01(w[0], w[1])
    u2 = u01(w[2], w[3])
    r = math.sqrt(-2.0 * math.log(u1))
    return r * math.cos(2.0 * math.pi * u2), r * math.sin(2.0 * math.pi * u2)
