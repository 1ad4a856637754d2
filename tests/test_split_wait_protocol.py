"""A model of the split-job rendezvous waits of the persistent batch engine.

A simulation split over P blocks is P consecutive queue entries ("parts").
Resident blocks claim entries in queue order as they become free
(`k_batch` / `serve_claim`, cave_amd/csrc/batch.hip).  Each part runs a phase
(the initial cells, then each tick), arrives at the job's rendezvous and waits
until all P parts have arrived (`part_rendezvous`, `part_rendezvous_tick`).  A
wait is bounded, to turn a hung job into an error instead of a hung device.

The bound must not fire on a healthy batch.  A part may wait a long time for
partners that have not been claimed yet, because every other block is busy
with earlier entries; that wait lasts as long as the work ahead of it, which
is unbounded in a large batch.  The kernel's rule therefore starts a part's
wait clock only once every part of its job has been claimed (the claim
counter has passed the job's last entry); from then on a wait lasts at most
one phase of a partner.  This model runs random batches in discrete time and
checks that, under the kernel's rule, no measured wait exceeds the longest
phase, and that every job completes when no job has more parts than a quarter
of the resident blocks (the planner's cap).  The earlier rule (the clock
starts at arrival) is the negative control: some batch must exceed the bound
under it, which shows the model can catch the problem at all.
"""
import random

import pytest


def run_batch(rng, blocks, jobs, max_phase):
    """jobs: list of (parts, phases).  Returns (max wait measured from
    arrival, max wait measured from max(arrival, every part claimed), all
    jobs finished)."""
    queue = [(j, p) for j, (parts, _) in enumerate(jobs) for p in range(parts)]
    nxt = 0
    claimed = [0] * len(jobs)
    all_claimed_at = [None] * len(jobs)
    block = [None] * blocks  # (job, part) or None
    remaining = {}           # (job, part) -> time left in its current phase; None while waiting
    phase = [0] * len(jobs)
    arrived = [set() for _ in jobs]
    arrival_time = {}
    done = [False] * len(jobs)
    max_old = max_new = 0
    t = 0
    limit = 10 * sum(parts * phases * max_phase for parts, phases in jobs) + 100
    while not all(done) and t < limit:
        # free blocks claim the next entries in queue order
        for b in range(blocks):
            if block[b] is None and nxt < len(queue):
                j, p = queue[nxt]
                nxt += 1
                block[b] = (j, p)
                claimed[j] += 1
                if claimed[j] == jobs[j][0]:
                    all_claimed_at[j] = t
                remaining[(j, p)] = rng.randint(1, max_phase)
        # one unit of time
        t += 1
        for b in range(blocks):
            if block[b] is None:
                continue
            j, p = block[b]
            if remaining[(j, p)] is not None:
                remaining[(j, p)] -= 1
                if remaining[(j, p)] == 0:
                    remaining[(j, p)] = None
                    arrived[j].add(p)
                    arrival_time[(j, p)] = t
        # rendezvous: a job whose parts have all arrived releases them
        for j, (parts, phases) in enumerate(jobs):
            if done[j] or len(arrived[j]) < parts:
                continue
            for p in range(parts):
                a = arrival_time[(j, p)]
                max_old = max(max_old, t - a)
                max_new = max(max_new, t - max(a, all_claimed_at[j]))
            arrived[j].clear()
            phase[j] += 1
            if phase[j] == phases:
                done[j] = True
                for b in range(blocks):
                    if block[b] is not None and block[b][0] == j:
                        block[b] = None
            else:
                for p in range(parts):
                    remaining[(j, p)] = rng.randint(1, max_phase)
    return max_old, max_new, all(done)


def random_batch(rng, blocks):
    jobs = []
    for _ in range(rng.randint(5, 40)):
        parts = 1 if rng.random() < 0.5 else rng.randint(2, max(2, blocks // 4))
        jobs.append((parts, rng.randint(1, 12)))
    return jobs


def test_wait_clock_starts_once_every_part_is_claimed():
    rng = random.Random(2024)
    max_phase = 5
    old_exceeded = 0
    for trial in range(400):
        blocks = rng.choice([8, 12, 16, 24])
        jobs = random_batch(rng, blocks)
        max_old, max_new, finished = run_batch(rng, blocks, jobs, max_phase)
        assert finished, f"trial {trial}: a job never completed"
        # kernel rule: a wait is at most one partner phase
        assert max_new <= max_phase, f"trial {trial}: wait {max_new} > {max_phase}"
        old_exceeded += max_old > max_phase
    # negative control: measured from arrival, waits for unclaimed partners exceed the bound
    assert old_exceeded > 0


def test_more_parts_than_blocks_cannot_finish():
    """Why the planner caps the parts of a job: with more parts than
    resident blocks, the claimed parts wait for partners no block can take."""
    rng = random.Random(7)
    _, _, finished = run_batch(rng, 4, [(1, 3), (6, 2)], 3)
    assert not finished


@pytest.mark.parametrize("blocks", [8, 16])
def test_jobs_at_the_parts_cap_finish(blocks):
    rng = random.Random(blocks)
    jobs = [(blocks // 4, 10)] * 12 + [(1, 30)] * 8
    rng.shuffle(jobs)
    _, max_new, finished = run_batch(rng, blocks, jobs, 4)
    assert finished and max_new <= 4
