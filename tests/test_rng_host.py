"""Host-side PCG64 jumps (the rewind a batch applies when it hands an
instance on with its next noise already drawn), against NumPy's PCG64."""

import numpy as np

from hit_solver_amd.rng import pcg64_advance, pcg64_words, pcg64_words_rewind


def _state(bg):
    s = bg.state["state"]
    return int(s["state"]), int(s["inc"])


def test_pcg64_advance_matches_numpy():
    for seed, k in [(0, 1), (1235, 262144 + 3017), (63, 10**9 + 7)]:
        bg = np.random.PCG64(seed)
        s0, inc = _state(bg)
        bg.advance(k)
        assert pcg64_advance(s0, inc, k) == _state(bg)[0]


def test_pcg64_rewind_undoes_a_draw():
    """Rewinding by the words a draw consumed restores the state, and the
    same normals come out again."""
    g = np.random.default_rng(4242)
    w0 = pcg64_words(g.bit_generator)
    ref = g.standard_normal(5000)
    # words consumed: compare with raw-word advances of a copy
    h = np.random.PCG64()
    h.state = np.random.default_rng(4242).bit_generator.state
    target = _state(g.bit_generator)[0]
    k = 0
    while _state(h)[0] != target:
        h.random_raw(1)
        k += 1
    assert k >= 5000
    w1 = pcg64_words(g.bit_generator)
    back = pcg64_words_rewind(w1, k)
    np.testing.assert_array_equal(back, w0)
    g2 = np.random.Generator(np.random.PCG64())
    st = g2.bit_generator.state
    st["state"]["state"] = (int(back[0]) << 64) | int(back[1])
    st["state"]["inc"] = (int(back[2]) << 64) | int(back[3])
    g2.bit_generator.state = st
    np.testing.assert_array_equal(g2.standard_normal(5000), ref)
