"""A model of the resident serving kernel's claim / close protocol.

`serve_claim` (cave_amd/csrc/batch.hip) runs in thread 0 of every block of
the persistent kernel.  Blocks claim ring entries in order; the block that
holds the poll lock reads the host's tail, raises `tail_seen` and may close
the kernel.  A request split into P parts is published as P consecutive
entries, and its parts wait for one another, so a kernel that closes while a
request is only partly claimed leaves the claimed parts waiting for partners
that never start, and cannot end.

The model below runs the protocol as a sequence of atomic steps (one shared
read, write, CAS or fetch-max each) for several blocks and a host that
publishes whole requests, in random interleavings, and checks that no request
is ever left partly claimed when every block has left the claim loop.  The
kernel's rule (closing decided under the lock, after a closing check, against
the largest tail seen) must pass every schedule; the earlier rule (no closing
check under the lock, closing against the holder's own tail read) is kept as
a negative control and must fail some schedule, which shows the check can
catch the problem at all.  Idle time is a free choice of the scheduler.
"""
import random

import pytest


class Shared:
    def __init__(self):
        self.head = 0
        self.tail_seen = 0
        self.closing = 0
        self.lock = 0
        self.tail = 0  # host memory


def block_agent(sh, claims, fixed, rng):
    """One block's claim loop; yields after every shared access.  Returns
    when the block leaves the loop (the kernel closed for it)."""
    while True:
        yield
        if sh.closing:
            return
        yield
        h = sh.head
        yield
        seen = sh.tail_seen
        if h < seen:
            yield
            if sh.head == h:  # CAS head h -> h + 1
                sh.head = h + 1
                claims.append(h)
            continue
        yield
        if sh.lock:  # CAS lock 0 -> 1 failed: nap, retry
            continue
        sh.lock = 1
        if fixed:
            yield
            if sh.closing:
                sh.lock = 0
                return
        yield
        t = sh.tail
        yield
        before = sh.tail_seen
        sh.tail_seen = max(before, t)
        limit = max(before, t) if fixed else t
        yield
        now_head = sh.head
        close = now_head >= limit and rng.random() < 0.5  # idle long enough: scheduler's choice
        if close:
            yield
            sh.closing = 1
        yield
        sh.lock = 0
        if close:
            return


def host_agent(sh, requests, rng, n_requests):
    """Publishes whole requests (1..4 parts each) by raising the tail."""
    for _ in range(n_requests):
        for _ in range(rng.randint(0, 6)):
            yield
        parts = rng.randint(1, 4)
        requests.append((sh.tail, parts))
        sh.tail += parts
        yield


def run_schedule(seed, fixed, n_blocks=3, n_requests=6):
    """Runs one random interleaving until every block left the loop (or a
    step bound).  Returns the requests left partly claimed."""
    rng = random.Random(seed)
    sh = Shared()
    claims, requests = [], []
    agents = [block_agent(sh, claims, fixed, rng) for _ in range(n_blocks)]
    host = host_agent(sh, requests, rng, n_requests)
    live = list(range(n_blocks))
    host_live = True
    for _ in range(20000):
        if not live:
            break
        pick = rng.randrange(len(live) + (1 if host_live else 0))
        if pick == len(live):
            try:
                next(host)
            except StopIteration:
                host_live = False
            continue
        try:
            next(agents[live[pick]])
        except StopIteration:
            live.pop(pick)
    if live:
        return []  # still running at the bound: nothing stranded yet
    claimed = set(claims)
    return [(s, p) for (s, p) in requests
            if any(s + k in claimed for k in range(p)) and not all(s + k in claimed for k in range(p))]


def test_closing_rule_never_strands_a_split_request():
    for seed in range(3000):
        assert run_schedule(seed, fixed=True) == [], seed


def test_model_catches_the_earlier_closing_rule():
    """Negative control: the rule without the closing check under the lock
    and against the holder's own tail read strands a request in some
    schedule."""
    stranded = [seed for seed in range(3000) if run_schedule(seed, fixed=False)]
    assert stranded, "the model found no stranding under the earlier rule"


@pytest.mark.parametrize("blocks", [2, 5])
def test_every_published_entry_is_claimed_once(blocks):
    """Claims are distinct and in ring order whatever the schedule."""
    for seed in range(300):
        rng = random.Random(seed)
        sh = Shared()
        claims, requests = [], []
        agents = [block_agent(sh, claims, True, rng) for _ in range(blocks)]
        host = host_agent(sh, requests, rng, 5)
        live, host_live = list(range(blocks)), True
        for _ in range(20000):
            if not live:
                break
            pick = rng.randrange(len(live) + (1 if host_live else 0))
            if pick == len(live):
                try:
                    next(host)
                except StopIteration:
                    host_live = False
                continue
            try:
                next(agents[live[pick]])
            except StopIteration:
                live.pop(pick)
        assert claims == sorted(set(claims))
        assert all(c < sh.tail for c in claims)
