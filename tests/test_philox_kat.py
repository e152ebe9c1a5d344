"""The counter-based generator of every device draw, philox4x32_10
(din-mol-li_amd/csrc/dml_internal.h, compiled for host and device from the
same source), against the three philox4x32-10 known-answer vectors of
Salmon et al., SC11 (Random123 kat_vectors)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = os.path.join(ROOT, "din-mol-li_amd", "dml", "philox_kat")

EXPECTED = ["6627e8d5 e169c58d bc57ac4c 9b00dbd8",
            "408f276d 41c83b0e a20bc7c6 6d5451fd",
            "d16cfe09 94fdcceb 5001e420 24126ea1"]


def test_philox4x32_10_known_answers():
    out = subprocess.run([KAT], capture_output=True, text=True, timeout=30)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split("\n")[:3] == EXPECTED


def test_host_restatement_known_answers():
    """tests/philox_host.py (the checker of the device streams) on the same vectors"""
    import philox_host as ph
    vec = [((0, 0, 0, 0), (0, 0)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0))]
    got = [" ".join("%08x" % w for w in ph.philox4x32_10(c, k)) for c, k in vec]
    assert got == EXPECTED
