"""ctypes binding of the C ABI in include/hss.h (libhss.so beside this file).

The library is the product's only compute path: there is no CPU fallback, and
importing this module raises when libhss.so has not been built
(``make -C heronian-square-search_amd``).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libhss.so"

HSS_OK, HSS_EINVAL, HSS_ENOMEM, HSS_EDEVICE, HSS_EOVERFLOW, HSS_EAGAIN = 0, 1, 2, 3, 4, 5
REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_FAILED, REQ_CANCELED = 0, 1, 2, 3, 4


class Hit(ctypes.Structure):
    _fields_ = [("l", ctypes.c_uint64), ("m", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("area4_lo", ctypes.c_uint64), ("area4_hi", ctypes.c_uint64)]

    @property
    def area(self) -> int:
        return ((self.area4_hi << 64) | self.area4_lo) // 4

    def line(self) -> str:
        return f"{self.l} {self.m} {self.n} {self.area}"


class RootResult(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint64), ("leaves", ctypes.c_uint64), ("pairs", ctypes.c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [("plan_ms", ctypes.c_double), ("setup_ms", ctypes.c_double), ("cells_ms", ctypes.c_double),
                ("exact_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("tiles", ctypes.c_uint64),
                ("left_entries", ctypes.c_uint64), ("deferred", ctypes.c_uint64), ("raw_roots", ctypes.c_uint64),
                ("leaves", ctypes.c_uint64), ("candidates", ctypes.c_uint64), ("grid_blocks", ctypes.c_uint32),
                ("block_threads", ctypes.c_uint32), ("window_tiles", ctypes.c_uint64),
                ("sorted_lefts", ctypes.c_uint64), ("evaluated", ctypes.c_uint64), ("window_evaluated", ctypes.c_uint64),
                ("window_ms", ctypes.c_double),
                ("plain_ms", ctypes.c_double), ("graph_launch", ctypes.c_uint64),
                ("window_candidates", ctypes.c_uint64), ("plain_candidates", ctypes.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


WEIGHT_CLASSES = 9


class Tuning(ctypes.Structure):
    _fields_ = [("tiles_per_wave", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("win_max_frac", ctypes.c_double * WEIGHT_CLASSES), ("win_item_cost", ctypes.c_double),
                ("win_partner_cost", ctypes.c_double), ("plain_partner_cost", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {"tiles_per_wave": self.tiles_per_wave, "win_max_frac": list(self.win_max_frac),
                "win_item_cost": self.win_item_cost, "win_partner_cost": self.win_partner_cost,
                "plain_partner_cost": self.plain_partner_cost}

    @classmethod
    def from_dict(cls, d: dict) -> "Tuning":
        t = cls()
        t.tiles_per_wave = d["tiles_per_wave"]
        for i, v in enumerate(d["win_max_frac"]):
            t.win_max_frac[i] = v
        t.win_item_cost = d["win_item_cost"]
        t.win_partner_cost = d["win_partner_cost"]
        t.plain_partner_cost = d.get("plain_partner_cost", 0.0)
        return t


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_slots", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("state_bytes", ctypes.c_uint64),
                ("launch_work", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("limit", ctypes.c_uint64), ("first", ctypes.c_uint64),
                ("last", ctypes.c_uint64), ("resume", ctypes.c_uint64), ("candidates", ctypes.c_uint64),
                ("leaves", ctypes.c_uint64), ("pairs", ctypes.c_uint64), ("n_hits", ctypes.c_uint32),
                ("launches", ctypes.c_uint32), ("state", ctypes.c_int32), ("error", ctypes.c_int32),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("launches", ctypes.c_uint64), ("reruns", ctypes.c_uint64), ("shared_launches", ctypes.c_uint64),
                ("positions", ctypes.c_uint64),
                ("leaves", ctypes.c_uint64), ("candidates", ctypes.c_uint64), ("gpu_ms", ctypes.c_double),
                ("uptime_ms", ctypes.c_double), ("queued", ctypes.c_uint32), ("active", ctypes.c_uint32),
                ("max_active", ctypes.c_uint32), ("pool_positions", ctypes.c_uint32), ("pool_free", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


def load_library(path: Path = LIB_PATH) -> ctypes.CDLL:
    if not path.is_file():
        raise RuntimeError(f"{path} is missing: build it with `make -C {HERE}` (no CPU fallback exists)")
    lib = ctypes.CDLL(str(path))
    u64, u32, vp = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p
    lib.hss_ctx_create.argtypes = [u64, ctypes.c_int, ctypes.POINTER(vp)]
    lib.hss_ctx_destroy.argtypes = [vp]
    lib.hss_ctx_destroy.restype = None
    lib.hss_last_error.restype = ctypes.c_char_p
    lib.hss_ctx_nmax.argtypes = [vp]
    lib.hss_ctx_nmax.restype = u64
    lib.hss_ctx_device_bytes.argtypes = [vp]
    lib.hss_ctx_device_bytes.restype = u64
    lib.hss_search.argtypes = [vp, u64, ctypes.POINTER(u64), u32, ctypes.POINTER(RootResult),
                               ctypes.POINTER(u64), ctypes.POINTER(Hit), u32, ctypes.POINTER(u32)]
    lib.hss_search_interval.argtypes = [vp, u64, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(Hit), u32,
                                        ctypes.POINTER(u32)]
    lib.hss_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    lib.hss_ctx_use_graphs.argtypes = [vp, ctypes.c_int]
    lib.hss_ctx_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.hss_ctx_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.hss_autotune.argtypes = [vp, u64, ctypes.POINTER(u64), u32, u32, ctypes.POINTER(Tuning),
                                 ctypes.POINTER(ctypes.c_double)]
    lib.hss_verify_triple.argtypes = [u64, u64, u64, u64, u64]
    lib.hss_estimate.argtypes = [vp, u64, u64, u64, ctypes.POINTER(ctypes.c_double)]
    lib.hss_split_by_work.argtypes = [vp, u64, u64, u64, u32, ctypes.POINTER(u64)]
    lib.hss_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    lib.hss_server_submit.argtypes = [vp, u64, u64, u64, ctypes.POINTER(u64)]
    lib.hss_server_wait.argtypes = [vp, u64, ctypes.c_int64, ctypes.POINTER(RequestInfo)]
    lib.hss_server_results.argtypes = [vp, u64, ctypes.POINTER(RootResult), u64, ctypes.POINTER(Hit), u32,
                                       ctypes.POINTER(u32)]
    lib.hss_server_release.argtypes = [vp, u64]
    lib.hss_server_cancel.argtypes = [vp, u64]
    lib.hss_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
    lib.hss_server_destroy.argtypes = [vp]
    lib.hss_server_destroy.restype = None
    return lib


_lib = load_library()


class HssError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hss error {code}: {msg}")
        self.code = code


def _check(rc: int):
    if rc != HSS_OK:
        raise HssError(rc, _lib.hss_last_error().decode(errors="replace"))


def verify_triple(l: int, m: int, n: int, area: int) -> bool:
    """Exact Heron check of a result line (the counterpart of scripts/verify.py)."""
    return bool(_lib.hss_verify_triple(l, m, n, area & ((1 << 64) - 1), area >> 64))


@dataclass
class SearchResult:
    candidates: list[int]
    leaves: list[int]
    pairs: list[int]
    hits: list[tuple[int, int, int, int]]
    stats: dict

    @property
    def total(self) -> int:
        return sum(self.candidates)


class Context:
    """Device-resident tables for every LIMIT <= n_max on one HIP device."""

    def __init__(self, n_max: int, device: int = 0):
        self._h = ctypes.c_void_p()
        _check(_lib.hss_ctx_create(n_max, device, ctypes.byref(self._h)))
        self.n_max = n_max
        self._server = None

    def close(self):
        if self._server is not None:
            self._server.close()
        if self._h:
            _lib.hss_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_bytes(self) -> int:
        return _lib.hss_ctx_device_bytes(self._h)

    def stats(self) -> dict:
        s = Stats()
        _check(_lib.hss_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def use_graphs(self, enable: bool) -> bool:
        """Replay the kernel chain as a captured hipGraph (default) or launch it
        directly; returns the previous setting."""
        return bool(_lib.hss_ctx_use_graphs(self._h, 1 if enable else 0))

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(_lib.hss_ctx_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw) -> dict:
        """Change scheduling knobs (counts never depend on them); returns the old
        ones.  win_max_frac may be a list (one per weight class) or one number
        for every class."""
        old = self.get_tuning()
        new = dict(old, **kw)
        if not isinstance(new["win_max_frac"], (list, tuple)):
            new["win_max_frac"] = [float(new["win_max_frac"])] * WEIGHT_CLASSES
        t = Tuning.from_dict(new)
        _check(_lib.hss_ctx_set_tuning(self._h, ctypes.byref(t)))
        return old

    def autotune(self, limit: int, roots, repeats: int = 1) -> tuple[dict, float]:
        """Run the autotuner on calibration roots; applies and returns the best
        tuning and its device time (ms)."""
        roots = [int(r) for r in roots]
        arr = (ctypes.c_uint64 * len(roots))(*roots)
        t = Tuning()
        ms = ctypes.c_double()
        _check(_lib.hss_autotune(self._h, limit, arr, len(roots), repeats, ctypes.byref(t), ctypes.byref(ms)))
        return t.as_dict(), ms.value

    def search(self, limit: int, roots, hit_cap: int = 4096) -> SearchResult:
        roots = [int(r) for r in roots]
        n = len(roots)
        arr = (ctypes.c_uint64 * max(n, 1))(*roots)
        per = (RootResult * max(n, 1))()
        total = ctypes.c_uint64()
        hits = (Hit * max(hit_cap, 1))()
        nh = ctypes.c_uint32()
        _check(_lib.hss_search(self._h, limit, arr, n, per, ctypes.byref(total), hits, hit_cap, ctypes.byref(nh)))
        return SearchResult([per[i].candidates for i in range(n)], [per[i].leaves for i in range(n)],
                            [per[i].pairs for i in range(n)],
                            [(hits[i].l, hits[i].m, hits[i].n, hits[i].area) for i in range(nh.value)],
                            self.stats())

    def search_interval(self, limit: int, first: int, last: int, hit_cap: int = 4096):
        total = ctypes.c_uint64()
        hits = (Hit * max(hit_cap, 1))()
        nh = ctypes.c_uint32()
        _check(_lib.hss_search_interval(self._h, limit, first, last, ctypes.byref(total), hits, hit_cap,
                                        ctypes.byref(nh)))
        return total.value, [(hits[i].l, hits[i].m, hits[i].n, hits[i].area) for i in range(nh.value)]

    def estimate(self, limit: int, first: int, last: int) -> float:
        """Planner work estimate (cells) of the positions [first, last] at limit."""
        out = ctypes.c_double()
        _check(_lib.hss_estimate(self._h, limit, first, last, ctypes.byref(out)))
        return out.value

    def split_by_work(self, limit: int, first: int, last: int, parts: int) -> list[int]:
        """Bounds of `parts` contiguous intervals of [first, last] with about
        equal estimated work: interval i is [b[i], b[i+1] - 1]."""
        out = (ctypes.c_uint64 * (parts + 1))()
        _check(_lib.hss_split_by_work(self._h, limit, first, last, parts, out))
        return list(out)


@dataclass
class ServedResult:
    info: dict
    candidates: list[int]
    leaves: list[int]
    pairs: list[int]
    hits: list[tuple[int, int, int, int]]

    @property
    def total(self) -> int:
        return sum(self.candidates)


class Server:
    """The served form over one Context: submit interval requests from any
    thread, wait for them; requests share launches on the device (continuous
    batching) and are admitted by request slots and device state memory.
    The context cannot run Context.search while a server is attached."""

    def __init__(self, ctx: Context, max_slots: int = 0, state_bytes: int = 0, launch_work: float = 0.0):
        self._ctx = ctx
        self._h = ctypes.c_void_p()
        cfg = ServerConfig(max_slots, 0, state_bytes, launch_work)
        _check(_lib.hss_server_create(ctx._h, ctypes.byref(cfg), ctypes.byref(self._h)))
        ctx._server = self

    def submit(self, limit: int, first: int, last: int) -> int:
        rid = ctypes.c_uint64()
        _check(_lib.hss_server_submit(self._h, limit, first, last, ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1) -> dict | None:
        """Request info once done; None when still running at the timeout."""
        info = RequestInfo()
        rc = _lib.hss_server_wait(self._h, rid, timeout_ms, ctypes.byref(info))
        if rc == HSS_EAGAIN:
            return None
        _check(rc)
        return info.as_dict()

    def info(self, rid: int) -> dict:
        info = RequestInfo()
        rc = _lib.hss_server_wait(self._h, rid, 0, ctypes.byref(info))
        if rc not in (HSS_OK, HSS_EAGAIN):
            _check(rc)
        return info.as_dict()

    def result(self, rid: int, hit_cap: int = 4096) -> ServedResult:
        info = self.wait(rid)
        if info is None or info["state"] == REQ_FAILED:
            raise HssError(HSS_EDEVICE, "request failed")
        n = info["last"] - info["first"] + 1
        per = (RootResult * n)()
        hits = (Hit * max(hit_cap, 1))()
        nh = ctypes.c_uint32()
        _check(_lib.hss_server_results(self._h, rid, per, n, hits, hit_cap, ctypes.byref(nh)))
        return ServedResult(info, [per[i].candidates for i in range(n)], [per[i].leaves for i in range(n)],
                            [per[i].pairs for i in range(n)],
                            [(hits[i].l, hits[i].m, hits[i].n, hits[i].area) for i in range(nh.value)])

    def release(self, rid: int):
        _check(_lib.hss_server_release(self._h, rid))

    def cancel(self, rid: int):
        """Stop a request; it ends REQ_CANCELED with resume at its finished prefix."""
        _check(_lib.hss_server_cancel(self._h, rid))

    def partial(self, rid: int, hit_cap: int = 4096) -> ServedResult:
        """Results so far (finished positions and hits found so far), without waiting."""
        info = self.info(rid)
        n = info["last"] - info["first"] + 1
        per = (RootResult * n)()
        hits = (Hit * max(hit_cap, 1))()
        nh = ctypes.c_uint32()
        _check(_lib.hss_server_results(self._h, rid, per, n, hits, hit_cap, ctypes.byref(nh)))
        return ServedResult(info, [per[i].candidates for i in range(n)], [per[i].leaves for i in range(n)],
                            [per[i].pairs for i in range(n)],
                            [(hits[i].l, hits[i].m, hits[i].n, hits[i].area) for i in range(nh.value)])

    def stats(self) -> dict:
        s = ServerStats()
        _check(_lib.hss_server_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        """Finish every submitted request and detach from the context."""
        if self._h:
            _lib.hss_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            if self._ctx._server is self:
                self._ctx._server = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def visible_device() -> int:
    return int(os.environ.get("HSS_DEVICE", "0"))
