"""Trigger sensors: enter / stay / exit events from the device overlap set.

A trigger detects overlap but generates no impulse (reference
``triggers.py:1-8``).  The device finds, each step, the broadphase pairs with a
trigger body whose shapes intersect after the contact solve (``k_trigger``,
read back by ``omi_batch_read_trigger_overlaps``); :class:`TriggerSystem`
diffs that set against the previous step's, as ``TriggerSystem.update``
(``triggers.py:24-45``) does, and the world dispatches the events to its
listeners.
"""
from typing import Iterable, List, Set, Tuple


class TriggerSystem:
    """Tracks sensor overlaps across steps and emits enter/stay/exit events."""

    def __init__(self) -> None:
        self.overlaps: Set[Tuple[int, int]] = set()   # {(trigger_body, other_body)}

    def update(self, overlaps: Iterable[Tuple[int, int]]) -> List[Tuple[str, int, int]]:
        """Diff this step's overlap pairs against the last; return ``(event, trigger_body,
        other_body)`` events: every ``enter``, then every ``stay``, then every ``exit``
        (each group in ascending body order; the reference iterates Python sets)."""
        new = {(int(t), int(o)) for t, o in overlaps}
        events = [('enter', t, o) for (t, o) in sorted(new - self.overlaps)]
        events += [('stay', t, o) for (t, o) in sorted(new & self.overlaps)]
        events += [('exit', t, o) for (t, o) in sorted(self.overlaps - new)]
        self.overlaps = new
        return events
