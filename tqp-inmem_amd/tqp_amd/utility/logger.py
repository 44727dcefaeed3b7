"""Timers and message logging with the call surface the operators use.

The operators call `perf_logger().start(tag)` / `.stop(tag)` around every operator and pipeline
(utility/logger.py:198-227 of the reference); this module keeps that surface, plus `time()` as a
context manager and `report()` to a JSON file.  Unlike the reference, the loggers exist by default
(a disabled tracker costs one branch), so operators can be called without set-up.
"""
import json
import logging
import time
from collections import defaultdict
from contextlib import contextmanager


class PerformanceTracker:
    def __init__(self, logger: logging.Logger = None, enable: bool = False):
        self.logger = logger or logging.getLogger("tqp_amd.perf")
        self._stats = defaultdict(lambda: {'count': 0, 'total': 0.0})
        self._start_times = {}
        self._enable = enable

    def enable(self, on: bool = True):
        self._enable = on

    def start(self, tag: str):
        if not self._enable:
            return
        if tag in self._start_times:
            raise RuntimeError(f"Timing for '{tag}' already started.")
        self._start_times[tag] = time.perf_counter()

    def stop(self, tag: str):
        if not self._enable:
            return None
        if tag not in self._start_times:
            raise RuntimeError(f"Timing for '{tag}' was not started.")
        elapsed = time.perf_counter() - self._start_times.pop(tag)
        stat = self._stats[tag]
        stat['count'] += 1
        stat['total'] += elapsed
        return elapsed

    @contextmanager
    def time(self, tag: str):
        self.start(tag)
        try:
            yield
        finally:
            self.stop(tag)

    def __getitem__(self, tag: str):
        stat = self._stats.get(tag, {'count': 0, 'total': 0.0})
        c, t = stat['count'], stat['total']
        return {'count': c, 'total': t, 'avg': t / c if c else 0.0}

    def report(self, filepath: str = "performance_stats.json") -> dict:
        data = {tag: self[tag] for tag in self._stats}
        with open(filepath, 'w') as f:
            json.dump(data, f, indent=4)
        return data


_perf = PerformanceTracker()
_msg = logging.getLogger("tqp_amd")


def perf_logger() -> PerformanceTracker:
    return _perf


def set_perf_logger(tracker: PerformanceTracker):
    global _perf
    _perf = tracker


def message_logger() -> logging.Logger:
    return _msg
