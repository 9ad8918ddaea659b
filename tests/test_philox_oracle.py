"""The Python Philox4x32-10 checker against the Random123 known-answer vectors."""
from oracle import philox


def test_random123_known_answers():
    M = 0xFFFFFFFF
    assert philox.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C,
                                                         0x9B00DBD8]
    assert philox.philox4x32_10([M] * 4, [M, M]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                     0x6D5451FD]
    assert philox.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB,
                                                              0x5001E420, 0x24126EA1]


def test_uniforms_are_in_open_interval():
    for r in range(50):
        u1, u2 = philox.uniforms(12345, r, 3)
        assert 0.0 < u1 < 1.0 and 0.0 < u2 < 1.0
