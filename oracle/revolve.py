"""Revolve (binomial checkpointing) schedule -- TEST INFRASTRUCTURE ONLY.

Plain-Python restatement of implexity-ad's ``BinomialSchedule`` (crates/implexity-ad/
src/revolve.rs:38-388), used by tests to check the product's C++ schedule action by
action.  Actions are tuples:

    ("advance", from, to) | ("snapshot", step, slot) | ("restore", step, slot)
    | ("reverse", step) | ("release", slot)

Pinned by: ``forward_steps == optimal_forward_steps(n, s)`` (the reference's own
debug assertion, revolve.rs:291), exhaustive replay on small n (every Reverse(k)
sees states k-1 and k), and the closed form ``beta(s, r) = C(s+r, r)``.
"""
from __future__ import annotations

SATURATED = 1 << 100


def beta(s: int, r: int) -> int:
    """revolve.rs:60-71 -- C(s+r, r), saturated."""
    if r < 0:
        return 0
    value = 1
    for i in range(1, r + 1):
        value = value * (s + i) // i
        if value >= SATURATED:
            return SATURATED
    return value


def repetitions(steps: int, snapshots: int) -> int:
    """revolve.rs:78-93."""
    if steps <= 1:
        return 0
    if snapshots == 0:
        raise ValueError("no snapshots")
    r, value = 0, 1
    while value < steps:
        r += 1
        value = value * (snapshots + r) // r
    return r


def chain_advances(length: int, snapshots: int) -> int:
    """revolve.rs:95-109."""
    if length <= 1:
        return 0
    r, value = 0, 1
    while value < length:
        r += 1
        value = value * (snapshots + r) // r
    return r * length - value * r // (snapshots + 1)


def optimal_forward_steps(steps: int, snapshots: int) -> int:
    """revolve.rs:117-125."""
    if steps == 0:
        return 0
    if steps == 1:
        return 1
    return chain_advances(steps, snapshots) + 1


def split(length: int, slots: int) -> int:
    """revolve.rs:237-253."""
    if slots == 0:
        raise ValueError("a chain of more than one step needs a snapshot slot")
    if slots == 1:
        return length - 1
    r = repetitions(length, slots)
    lo = max(1, beta(slots, r - 2), max(0, length - beta(slots - 1, r)))
    hi = min(length - 1, beta(slots, r - 1), max(0, length - beta(slots - 1, r - 1)))
    if lo > hi:
        raise ValueError(f"no optimal split for a chain of {length} steps with {slots} slots")
    return hi


class _Builder:
    """revolve.rs:137-235."""

    def __init__(self, steps: int, total_slots: int):
        self.steps = steps
        self.actions: list[tuple] = []
        self.free = list(range(total_slots))[::-1]
        self.cursor = None
        self.in_use = 0
        self.peak = 0
        self.advanced = 0
        self.reverse_start = None

    def take_slot(self) -> int:
        if not self.free:
            raise ValueError("checkpoint schedule ran out of snapshot slots")
        slot = self.free.pop()
        self.in_use += 1
        self.peak = max(self.peak, self.in_use)
        return slot

    def release(self, slot: int) -> None:
        self.actions.append(("release", slot))
        self.in_use = max(0, self.in_use - 1)
        self.free.append(slot)
        self.free.sort(reverse=True)

    def advance(self, a: int, b: int) -> None:
        if b > a:
            self.actions.append(("advance", a, b))
            self.advanced += b - a
            self.cursor = b

    def ensure_cursor(self, position: int, slot: int) -> None:
        if self.cursor != position:
            self.actions.append(("restore", position, slot))
            self.cursor = position

    def reverse_step(self, step: int) -> None:
        if self.reverse_start is None:
            self.actions.append(("advance", self.steps - 1, self.steps))
            self.advanced += 1
            self.reverse_start = len(self.actions)
        self.actions.append(("reverse", step))
        self.cursor = None

    def chain(self, a: int, length: int, own: int, slots: int) -> None:
        while True:
            if length == 0:
                return
            if length == 1:
                self.ensure_cursor(a, own)
                self.reverse_step(a + 1)
                return
            m = split(length, slots)
            self.ensure_cursor(a, own)
            self.advance(a, a + m)
            if length - m > 1:
                slot = self.take_slot()
                self.actions.append(("snapshot", a + m, slot))
                self.chain(a + m, length - m, slot, slots - 1)
                self.release(slot)
            else:
                self.reverse_step(a + length)
            length = m


def binomial_schedule(steps: int, snapshots: int):
    """BinomialSchedule::new(steps, snapshots, 0) (revolve.rs:261-295).

    Returns (actions, reverse_start, forward_steps, peak)."""
    if steps == 0:
        raise ValueError("a checkpoint schedule needs at least one step")
    if snapshots == 0:
        raise ValueError("a checkpoint schedule needs at least one snapshot slot")
    b = _Builder(steps, snapshots)
    own = b.take_slot()
    b.actions.append(("snapshot", 0, own))
    b.cursor = 0
    b.chain(0, steps, own, snapshots)
    return b.actions, (b.reverse_start or 0), b.advanced, b.peak


def replay_check(steps: int, snapshots: int) -> int:
    """Exhaustive replay: every Reverse(k) must hold state k-1 at the cursor and state k
    ahead (scan.rs:328-388 semantics).  Returns the number of forward step evaluations."""
    actions, _, _, _ = binomial_schedule(steps, snapshots)
    return replay_actions(steps, actions)


def replay_actions(steps: int, actions) -> int:
    """replay_check on a given action list (tuples as binomial_schedule builds them)."""
    slots = {}
    cursor, ahead = 0, None
    evaluations = 0
    expected_reverse = steps
    for act in actions:
        kind = act[0]
        if kind == "snapshot":
            assert cursor == act[1], act
            slots[act[2]] = act[1]
        elif kind == "restore":
            assert slots.get(act[2]) == act[1], act
            cursor = act[1]
        elif kind == "release":
            slots.pop(act[1], None)
        elif kind == "advance":
            assert cursor == act[1], act
            for t in range(act[1], act[2]):
                evaluations += 1
                if t + 1 == steps:
                    ahead = steps
                    break
                cursor = t + 1
        elif kind == "reverse":
            k = act[1]
            assert k == expected_reverse, (k, expected_reverse)
            assert cursor is not None and cursor + 1 == k and ahead == k, (act, cursor, ahead)
            ahead = cursor
            cursor = None
            expected_reverse -= 1
    assert expected_reverse == 0
    return evaluations
