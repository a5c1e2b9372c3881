"""Memory accounting of the serving loop, on the CPU: a batch reserves its
estimate, gives the reservation back when it ends (the workspaces stay as a
cache), and idle caches are freed only when another batch needs the room."""

from aospectrum_amd.serving import BandServer, _Resident


class _FakeNative:
    def __init__(self):
        self.trims = 0

    def trim(self):
        self.trims += 1
        return 0


def _server(budget):
    server = BandServer(memory_budget_bytes=budget)
    a, b = _Resident("a", _FakeNative()), _Resident("b", _FakeNative())
    server._bundles = {"a": a, "b": b}
    return server, a, b


def _end_batch(server, res):
    # what the worker does when a batch returns
    res.busy = False
    res.cached, res.reserved = res.reserved, 0


def test_reservation_released_after_batch():
    server, a, b = _server(100)
    assert server._reserve(a, 60)
    a.busy = True
    assert (a.reserved, a.cached) == (60, 0)
    # b cannot fit next to a running batch
    assert not server._reserve(b, 50)
    _end_batch(server, a)
    assert (a.reserved, a.cached) == (0, 60)
    assert server.memory_in_use() == {"a": (0, 60), "b": (0, 0)}
    # now b fits by freeing a's idle cache
    assert server._reserve(b, 50)
    assert a.native.trims == 1 and a.cached == 0
    assert (b.reserved, b.cached) == (50, 0)


def test_cache_kept_when_room_and_reused_by_same_bundle():
    server, a, b = _server(200)
    assert server._reserve(a, 80)
    _end_batch(server, a)
    assert server._reserve(b, 100)       # fits beside a's cache: nothing freed
    assert a.native.trims == 0 and a.cached == 80
    _end_batch(server, b)
    # a's next batch takes over its own cache: it holds the larger of the two
    assert server._reserve(a, 40)
    assert (a.reserved, a.cached) == (80, 0)
    assert a.native.trims == 0


def test_busy_bundle_is_never_trimmed():
    server, a, b = _server(100)
    assert server._reserve(a, 70)
    a.busy = True
    assert not server._reserve(b, 40)
    assert a.native.trims == 0 and a.reserved == 70
    _end_batch(server, a)
    assert server._reserve(b, 40) is True   # 70 cached + 40 > 100: a is trimmed
    assert a.native.trims == 1


def test_own_cache_larger_than_room_is_freed():
    server, a, b = _server(100)
    assert server._reserve(a, 90)
    _end_batch(server, a)
    assert server._reserve(b, 30)           # frees a's cache
    # b keeps 30 reserved; a's new batch of 60 fits, its (freed) cache is 0
    assert server._reserve(a, 60)
    assert (a.reserved, b.reserved) == (60, 30)


def test_admission_follows_arrival_across_bundles():
    """A bundle that has just finished a batch cannot take the memory again
    while an older batch of another bundle waits, unless both fit."""
    server, a, b = _server(100)
    assert server._reserve(a, 70)
    a.busy = True
    # b's batch (arrived at t=1) waits: it does not fit beside a's batch
    server._waiting["b"] = (1.0, 60)
    assert not server._admit(b, 60, 1.0)
    _end_batch(server, a)
    # a's next batch arrived later (t=2): it must not take the room b waits for
    server._waiting["a"] = (2.0, 70)
    assert not server._admit(a, 70, 2.0)
    # b goes first (a's idle cache is freed for it)
    assert server._admit(b, 60, 1.0)
    del server._waiting["b"]
    b.busy = True
    # a small later batch may start beside b when it leaves room
    assert server._admit(a, 30, 2.0)
