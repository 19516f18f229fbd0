"""The NumPy Philox4x32-10 (oracle/philox.py) against the known-answer vectors
published with the algorithm (Random123's kat_vectors for philox4x32 with 10
rounds), and its normals against N(0, 1)."""
import numpy as np
from scipy import stats

from oracle import philox

KAT = [  # counter, key, output
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_known_answers():
    for ctr, key, want in KAT:
        got = philox.philox4x32_10(np.array(ctr, np.uint32), np.array(key, np.uint32))
        assert [int(v) for v in got] == list(want)


def test_normals_are_standard():
    z = philox.normals4(12345, np.arange(20000), 3, 7).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02
    assert stats.kstest(z, "norm").pvalue > 1e-3
