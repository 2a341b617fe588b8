"""Host-side bookkeeping of the batched readjust level (no GPU): the arena
that holds a level's sequences and outputs, run on a CPU tensor."""
import numpy as np

from poy5_amd.readjust import _Arena


def test_arena_places_each_sequence_on_a_16_byte_boundary():
    rng = np.random.default_rng(1)
    hosts = [rng.integers(0, 32, rng.integers(0, 40)).astype(np.uint8) for _ in range(300)]
    hosts += [np.zeros(0, np.uint8), [1, 2, 3]]
    arena = _Arena(50000, "cpu")
    arena.take(5)
    placed = arena.put_all(hosts)
    t = arena.t.numpy()
    free = np.ones(len(t), bool)
    for h, (at, n) in zip(hosts, placed):
        assert at % 16 == 0 and n == len(h)
        assert np.array_equal(t[at:at + n], np.asarray(h, np.uint8))
        free[at:at + n] = False
    assert not t[free].any()
    flat, st = arena.fetch([a for a, _ in placed], [n for _, n in placed])
    for h, s in zip(hosts, st):
        assert np.array_equal(flat[s:s + len(h)], np.asarray(h, np.uint8))


def test_arena_refuses_what_does_not_fit():
    arena = _Arena(64, "cpu")
    arena.put_all([np.ones(40, np.uint8)])
    try:
        arena.put_all([np.ones(40, np.uint8)])
    except MemoryError:
        return
    raise AssertionError("the arena took more than it holds")
