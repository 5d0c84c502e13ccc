"""CPU model of the persistent kernels' device work queue (kernels.hip: wg_queue_items, queue_leave,
zstd_parse_queue), run under random interleavings of its atomic steps.

Each taker (a workgroup, or a wavefront in the parse) repeats: fetch i = next++ (at most n + 1
fetches, else error bit 0 and leave); if i < n take item i.  On leaving: taken += mine; fence;
d = done++; the taker that sees d == takers - 1 checks taken == n (else error bit 1) and sets next,
taken and done back to 0.  The model checks, for every schedule tried: every item is taken exactly
once, no taker waits for another (each one's steps are bounded whatever the others do), the counters
are back at zero when the launch ends (so launches can follow one another, as graph replays do, with
no reset), the trip bound is never reached, and a queue left with a stale counter is reported."""
import random

import pytest


class Queue:
    def __init__(self, n):
        self.n, self.next, self.done, self.taken, self.err = n, 0, 0, 0, 0


def taker(q, takers, log):
    """One taker as a generator: yields between atomic steps (any other taker may run there)."""
    n = q.n
    mine = 0
    it = 0
    while True:
        i = q.next
        q.next += 1  # atomicAdd(next, 1)
        yield
        if it > n:
            q.err |= 1
            break
        it += 1
        if i >= n:
            break
        log.append(i)  # body(i)
        mine += 1
        yield
    q.taken += mine  # atomicAdd(taken, mine)
    yield
    d = q.done
    q.done += 1  # atomicAdd(done, 1)
    yield
    if d == takers - 1:
        if q.taken != n:
            q.err |= 2
        q.next = 0
        yield
        q.taken = 0
        yield
        q.done = 0


def launch(q, takers, rng, late=0.3):
    """All takers of one launch, stepped in a random order; some start only after others ended
    (a grid larger than what is resident)."""
    log = []
    pending = list(range(takers))
    rng.shuffle(pending)
    running = []
    steps = 0
    while pending or running:
        if pending and (not running or rng.random() > late):
            running.append(taker(q, takers, log))
            pending.pop()
        g = rng.choice(running)
        try:
            next(g)
        except StopIteration:
            running.remove(g)
        steps += 1
        assert steps < 100 * (q.n + 2) * (takers + 1), "a taker did not finish"
    return log


@pytest.mark.parametrize("seed", range(40))
def test_every_item_once_and_counters_reset(seed):
    rng = random.Random(seed)
    n = rng.choice([0, 1, 2, 3, 7, 31, 64, 200])
    takers = rng.choice([1, 2, 3, 8, 33])
    q = Queue(n)
    for _ in range(3):  # replays of one captured pass: no reset between them
        log = launch(q, takers, rng)
        assert sorted(log) == list(range(n))
        assert (q.next, q.taken, q.done, q.err) == (0, 0, 0, 0)


def test_stale_counter_is_reported():
    rng = random.Random(1)
    q = Queue(10)
    q.next = 4  # items 0..3 would be skipped
    log = launch(q, 3, rng)
    assert sorted(log) == list(range(4, 10))
    assert q.err & 2
    assert (q.next, q.taken, q.done) == (0, 0, 0)  # the next launch starts clean


def test_takers_never_wait_for_each_other():
    # one taker alone runs to its end while the others have not even started: nothing it does
    # depends on them
    q = Queue(5)
    log = []
    g = taker(q, 4, log)
    for _ in g:
        pass
    assert log == [0, 1, 2, 3, 4] and q.done == 1 and q.err == 0
