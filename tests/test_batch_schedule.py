"""CPU emulation of a single-launch batch work-list walk: which item
orders can deadlock when one launch runs every pass of every plan and its
workgroups wait inside the kernel for the previous pass of their plan.

The library does not run such a launch: its batch runs one persistent
launch per pass level and kernel, with no waits inside a kernel, the order
between levels given by the launch boundaries.  This emulation records the
condition a single-launch design has to meet.

Items are the tiles of a batch's tasks in level order (every plan's pass 0,
then every pass 1, ...), item g of task t depending on every item of the
previous pass of the same plan.  Workgroups are dispatched in index order
into at most R resident slots; a resident workgroup processes one item per
step when that item's dependency is complete and waits otherwise; a
workgroup that has no item left exits and frees its slot.

* Static stride (workgroup b of a grid of G takes items b, b + G,
  b + 2G, ... and waits inside the launch for each item's dependency)
  deadlocks as soon as fewer than G workgroups are resident at once:
  resident workgroups wait for items owned by workgroups that cannot be
  dispatched until one of them exits.  It only progresses when the whole
  grid is co-resident, which nothing in a launch guarantees (other kernels
  and graph branches share the CUs; occupancy depends on the register
  count).
* Tickets (a workgroup takes the next item from an atomic counter after
  finishing its previous one) never wait on an item no running workgroup
  holds: every dependency has a smaller ticket, so the smallest unfinished
  ticket can always proceed, for any number of resident workgroups R >= 1.
"""
import numpy as np
import pytest


def make_batch(nplans, seed, passes=3):
    """tiles per pass for `nplans` plans (like the uneven configuration:
    tile counts vary per plan and pass)"""
    rng = np.random.default_rng(seed)
    tiles = rng.integers(1, 40, size=(nplans, passes))
    # level order: every plan's pass 0, then pass 1, ...
    tasks = []   # (plan, pass, ntiles)
    for p in range(passes):
        for q in range(nplans):
            tasks.append((q, p, int(tiles[q, p])))
    index = {(q, p): i for i, (q, p, _) in enumerate(tasks)}
    dep = [index[(q, p - 1)] if p > 0 else -1 for (q, p, _) in tasks]
    item_task = [i for i, (_, _, nt) in enumerate(tasks) for _ in range(nt)]
    return tasks, dep, item_task


def simulate(tasks, dep, item_task, R, scheme, G=None, max_steps=100000):
    """returns the number of steps to finish, or None on deadlock"""
    nitems = len(item_task)
    left = [nt for (_, _, nt) in tasks]   # tiles of each task not yet done
    done_items = 0
    if scheme == "static":
        queues = [list(range(b, nitems, G)) for b in range(G)]
        next_wg = 0
        resident = []
        steps = 0
        while done_items < nitems:
            while len(resident) < R and next_wg < G:
                resident.append(next_wg)
                next_wg += 1
            progressed = False
            finished = []
            for b in resident:
                if not queues[b]:
                    finished.append(b)
                    continue
                g = queues[b][0]
                d = dep[item_task[g]]
                if d < 0 or left[d] == 0:
                    queues[b].pop(0)
                    left[item_task[g]] -= 1
                    done_items += 1
                    progressed = True
                    if not queues[b]:
                        finished.append(b)
            for b in finished:
                resident.remove(b)
            if not progressed and not finished:
                return None
            steps += 1
            if steps > max_steps:
                return None
        return steps
    # tickets
    ticket = 0
    holding = [None] * R
    steps = 0
    while done_items < nitems:
        progressed = False
        for s in range(R):
            if holding[s] is None and ticket < nitems:
                holding[s] = ticket
                ticket += 1
            g = holding[s]
            if g is None:
                continue
            d = dep[item_task[g]]
            if d < 0 or left[d] == 0:
                left[item_task[g]] -= 1
                done_items += 1
                holding[s] = None
                progressed = True
        if not progressed:
            return None
        steps += 1
        if steps > max_steps:
            return None
    return steps


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_static_stride_deadlocks_without_full_coresidency(seed):
    tasks, dep, item_task = make_batch(24, seed)
    G = 64
    assert simulate(tasks, dep, item_task, R=G, scheme="static", G=G) is not None
    stuck = [R for R in (1, 2, 8, 16, 32, 48, 63)
             if simulate(tasks, dep, item_task, R=R, scheme="static", G=G) is None]
    assert stuck, "expected a deadlock with partial residency"
    assert 63 in stuck or 48 in stuck


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("R", [1, 2, 3, 7, 16, 64, 500])
def test_tickets_always_progress(seed, R):
    tasks, dep, item_task = make_batch(24, seed)
    steps = simulate(tasks, dep, item_task, R=R, scheme="tickets")
    assert steps is not None
    # never slower than one item per step with one workgroup
    assert steps <= len(item_task)


def test_planted_reversed_levels_deadlock_with_tickets():
    """the ticket argument needs every dependency to carry a smaller ticket:
    with the levels reversed (pass 2 first) one resident workgroup stalls,
    so the test can tell a broken item order from a correct one"""
    tasks, dep, item_task = make_batch(8, 5)
    nlev = 3
    per = len(tasks) // nlev
    order = [t for lev in reversed(range(nlev)) for t in range(lev * per, (lev + 1) * per)]
    rev_item_task = [t for t in order for _ in range(tasks[t][2])]
    assert simulate(tasks, dep, rev_item_task, R=1, scheme="tickets") is None
