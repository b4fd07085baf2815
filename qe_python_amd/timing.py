"""QE-style cumulative timers with the reference's interface
(``qepy_pw/timing.py:38-63``: ``measure(name)``, ``start()``,
``stop(name, started)``, ``snapshot()``, ``entries``).  The device path
synchronises the GPU at the end of each measured region, so wall times cover
the kernels they label."""

from __future__ import annotations

from dataclasses import dataclass
import time


@dataclass
class TimingEntry:
    cpu_seconds: float = 0.0
    wall_seconds: float = 0.0
    calls: int = 0


class _Region:
    __slots__ = ("registry", "name", "cpu0", "wall0")

    def __init__(self, registry, name):
        self.registry = registry
        self.name = name

    def __enter__(self):
        self.cpu0 = time.process_time()
        self.wall0 = time.perf_counter()

    def __exit__(self, *exc):
        self.registry.stop(self.name, (self.cpu0, self.wall0))


class TimingRegistry:
    def __init__(self) -> None:
        self.entries: dict[str, TimingEntry] = {}

    def measure(self, name: str) -> _Region:
        return _Region(self, name)

    def start(self) -> tuple[float, float]:
        return time.process_time(), time.perf_counter()

    def stop(self, name: str, started, calls: int = 1) -> None:
        cpu0, wall0 = started
        entry = self.entries.setdefault(name, TimingEntry())
        entry.cpu_seconds += time.process_time() - cpu0
        entry.wall_seconds += time.perf_counter() - wall0
        entry.calls += calls

    def snapshot(self) -> dict[str, TimingEntry]:
        return {k: TimingEntry(v.cpu_seconds, v.wall_seconds, v.calls)
                for k, v in self.entries.items()}
