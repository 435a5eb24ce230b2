"""The project's autotuner for launch shapes and frames per launch.

A workload is described by a name and a size class; its tunable choices are

  * the library's knobs (ssg_tune_set): the fused accumulate's tile rows (SSG_TUNE_ACC_ROWS) and
    target blocks per launch, which fixes the frames per accumulate chunk (SSG_TUNE_ACC_BLOCKS);
    the persistent ECC passes' blocks per CU (SSG_TUNE_ECC_GRID); the frames per sequence and
    chunk of the batched focus stack (SSG_TUNE_FOCUS_CHUNK);
  * caller-side choices the workload's `run` callable understands, such as the frames handed to
    one batched launch ("frames_per_launch").

`Autotuner.tune` times the candidates on the device by coordinate descent (each choice in turn,
the others held at their best so far; the median of `repeats` timed runs after one untimed run,
which also captures any graph), keeps the fastest, caches it in a JSON file keyed by device
(architecture and CU count), workload and size class, and applies it.  `Autotuner.lookup`
returns a cached choice; the shipped `tuning.json` beside this module holds the choices measured
on MI355X, and a user cache ($SSG_TUNE_CACHE, default ~/.cache/serstacker_amd/tune.json) takes
precedence over it.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import statistics
import time

from . import _lib

KNOBS = {"acc_rows": _lib.TUNE_ACC_ROWS, "acc_blocks": _lib.TUNE_ACC_BLOCKS, "ecc_grid": _lib.TUNE_ECC_GRID,
         "focus_chunk": _lib.TUNE_FOCUS_CHUNK, "pyr_strip": _lib.TUNE_PYR_STRIP}
SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def size_class(width: int, height: int, frames: int | None = None) -> str:
    """Size class of a frame geometry: the pixel count rounded to a power of two (in kpx), plus
    the batch size rounded to a power of two when given."""
    px = width * height
    k = 1
    while k * 1024 < px:
        k *= 2
    key = f"{k}kpx"
    if frames:
        b = 1
        while b < frames:
            b *= 2
        key += f"x{b}"
    return key


def device_name() -> str:
    n = C.c_int()
    buf = C.create_string_buffer(256)
    if _lib.load().ssg_device_info(C.byref(n), buf, 256) != 0:
        return "unknown"
    return buf.value.decode(errors="replace") or "unknown"


def device_key(name: str | None = None) -> str:
    """The cache key of a device: its architecture and CU count (the text in the parentheses of
    ssg_device_info's name), which the marketing name in front of it may not be stable about."""
    name = device_name() if name is None else name
    lo, hi = name.rfind("("), name.rfind(")")
    return name[lo + 1:hi] if 0 <= lo < hi else name


def get_knob(name: str) -> int:
    v = C.c_int32()
    _lib.check(_lib.load().ssg_tune_get(KNOBS[name], C.byref(v)), "ssg_tune_get")
    return v.value


def set_knob(name: str, value: int):
    _lib.check(_lib.load().ssg_tune_set(KNOBS[name], int(value)), "ssg_tune_set")


class Autotuner:
    def __init__(self, cache_path: str | None = None):
        self.path = cache_path or os.environ.get("SSG_TUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "serstacker_amd", "tune.json")
        self.device = device_key()

    @staticmethod
    def _read(path):
        try:
            with open(path) as fh:
                return json.load(fh)
        except (OSError, ValueError):
            return {}

    def _key(self, workload: str, size_key: str) -> str:
        return f"{self.device}|{workload}|{size_key}"

    def lookup(self, workload: str, size_key: str) -> dict | None:
        k = self._key(workload, size_key)
        for path in (self.path, SHIPPED):
            hit = self._read(path).get(k)
            if hit:
                return dict(hit["choice"])
        return None

    def store(self, workload: str, size_key: str, choice: dict, timings: dict):
        data = self._read(self.path)
        data[self._key(workload, size_key)] = {"choice": choice, "seconds": timings,
                                               "when": time.strftime("%Y-%m-%d %H:%M:%S")}
        os.makedirs(os.path.dirname(self.path) or ".", exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    @staticmethod
    def apply(choice: dict):
        """Sets the library knobs of a choice (caller-side entries are left to the caller)."""
        for k, v in choice.items():
            if k in KNOBS:
                set_knob(k, v)

    def tune(self, workload: str, size_key: str, space: dict, run, repeats: int = 2, retune: bool = False):
        """Returns (choice, timings, cached).  space: {choice name: [candidates]}, the first
        candidate of each the starting point; run(choice) performs one unit of the workload with
        the knobs already applied (the caller synchronises the device inside it)."""
        if not retune:
            hit = self.lookup(workload, size_key)
            if hit is not None and set(hit) == set(space):
                self.apply(hit)
                return hit, {}, True
        best = {k: v[0] for k, v in space.items()}
        timings = {}

        def measure(choice):
            key = json.dumps(choice, sort_keys=True)
            if key not in timings:
                self.apply(choice)
                run(choice)  # untimed: first use (graph capture, allocations)
                ts = []
                for _ in range(repeats):
                    t0 = time.perf_counter()
                    run(choice)
                    ts.append(time.perf_counter() - t0)
                timings[key] = statistics.median(ts)
            return timings[key]

        for name, cands in space.items():
            scored = []
            for c in cands:
                trial = dict(best)
                trial[name] = c
                scored.append((measure(trial), c))
            best[name] = min(scored)[1]
        self.apply(best)
        self.store(workload, size_key, best, timings)
        return best, timings, False
