"""Nested wall-clock timers with the interface of lib/global_clock.py (GlobalClock:
start / stop / report / top_level_clocks)."""
from __future__ import annotations

import time


class Clock:
    def __init__(self, name):
        self.name = name
        self.start = None
        self.tot_time = 0.0
        self.count = 0
        self.children = {}

    def stop(self):
        self.tot_time += time.perf_counter() - self.start
        self.count += 1


class GlobalClock:
    def __init__(self):
        self.top_level_clocks = {}
        self._stack = []

    def start(self, name):
        level = self._stack[-1].children if self._stack else self.top_level_clocks
        clk = level.get(name)
        if clk is None:
            clk = level[name] = Clock(name)
        clk.start = time.perf_counter()
        self._stack.append(clk)

    def stop(self, name):
        while self._stack:
            clk = self._stack.pop()
            clk.stop()
            if clk.name == name:
                return
        raise KeyError(f"clock {name} was not started")

    def report(self, file=None):
        def rec(clocks, depth):
            for c in clocks.values():
                print(f"{'  ' * depth}{c.name}: {c.tot_time:.6f} s ({c.count} calls)", file=file)
                rec(c.children, depth + 1)
        rec(self.top_level_clocks, 0)
