"""Python front end of the HIP engine (``bin/libgsamp_amd.so``, C ABI in include/gsamp_amd.h).

Mirrors the modes of the reference engine ``gsamp2`` (gsamp2.cpp:149-225) as methods of :class:`Engine`; bit
strings are numpy int8 arrays in STEP order ([nstr, ns]; -1 = '?').  There is no CPU fallback: constructing an
engine without the built library or without a HIP device raises.
"""
import ctypes
import hashlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "bin", "libgsamp_amd.so")
CLI_PATH = os.path.join(HERE, "bin", "gsamp2_amd")

C64, C128 = 0, 1
SEED_PLAIN, SEED_SEQ = 0, 1

# every symbol include/gsamp_amd.h declares
ABI_SYMBOLS = ("gs_plan_load", "gs_plan_parse", "gs_plan_info", "gs_plan_destroy", "gs_engine_create",
               "gs_engine_destroy", "gs_prob", "gs_chain", "gs_sample", "gs_hybrid", "gs_prefix", "gs_resume",
               "gs_last_device_ms", "gs_last_error", "gs_profile_enable", "gs_profile_read",
               "gs_step_geometry", "gs_step_columns", "gs_set_fusion", "gs_set_fused4", "gs_set_grid_log2", "gs_set_grid_log2_one_slot", "gs_set_small_steps", "gs_set_walk", "gs_set_xcd_blocks", "gs_set_plain_pairs", "gs_set_fused_run", "gs_set_diag_runs", "gs_step_run", "gs_set_fold_reduce", "gs_race_check", "gs_copy_probe", "gs_batch_create", "gs_batch_destroy", "gs_batch_score",
               "gs_batch_use_graph", "gs_batch_set_tile", "gs_batch_set_unroll", "gs_batch_set_xcd_items", "gs_batch_last_ms", "gs_batch_last_error",
               "gs_batch_set_matrices", "gs_batch_grad", "gs_batch_grad_ms", "gs_tt_plan_load", "gs_tt_plan_info",
               "gs_tt_plan_destroy", "gs_tt_engine_create", "gs_tt_engine_destroy", "gs_tt_prob", "gs_tt_chain",
               "gs_tt_sample", "gs_tt_hybrid", "gs_tt_prefix", "gs_tt_resume", "gs_tt_dp", "gs_tt_last_error", "gs_tt_set_grid_log2",
               "gs_prob_noalloc")

_lib = None


class GsError(RuntimeError):
    """A failed engine call; ``code`` is the reference's exit code for the same failure (see gsamp_amd.h)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def library():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or build.py) first")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, I64, D, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64
    pp = ctypes.POINTER(ctypes.c_void_p)
    i8p = ctypes.POINTER(ctypes.c_int8)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "gs_plan_load": (I, [ctypes.c_char_p, pp]),
        "gs_plan_parse": (I, [ctypes.c_char_p, pp]),
        "gs_plan_info": (I, [P, ip, ip, ip]),
        "gs_plan_destroy": (None, [P]),
        "gs_engine_create": (I, [P, I, I, pp]),
        "gs_engine_destroy": (None, [P]),
        "gs_prob": (I, [P, i8p, I64, dp]),
        "gs_prob_noalloc": (I, [P, I, I, i8p, I64, dp]),
        "gs_chain": (I, [P, i8p, I64, dp, dp, dp]),
        "gs_sample": (I, [P, U64, I64, I64, I, i8p, dp]),
        "gs_hybrid": (I, [P, i8p, I64, U64, I64, i8p, dp, dp]),
        "gs_prefix": (I, [P, i8p, I, dp, I64, ip, ip, dp, dp]),
        "gs_resume": (I, [P, i8p, I, I, ctypes.POINTER(dp), dp, ctypes.POINTER(U64), ctypes.POINTER(U64), U64, I64,
                          i8p, dp, dp, dp]),
        "gs_last_device_ms": (D, [P]),
        "gs_last_error": (ctypes.c_char_p, []),
        "gs_profile_enable": (I, [P, I]),
        "gs_profile_read": (I, [P, dp, ctypes.POINTER(I64)]),
        "gs_step_geometry": (I, [P, I, ip]),
        "gs_set_fusion": (I, [P, I]),
        "gs_set_fused4": (I, [P, I]),
        "gs_step_columns": (I, [P, I, I, ctypes.POINTER(ctypes.c_uint64), I]),
        "gs_set_grid_log2": (I, [P, I]),
        "gs_set_small_steps": (I, [P, I]),
        "gs_set_grid_log2_one_slot": (I, [P, I]),
        "gs_set_xcd_blocks": (I, [P, I]),
        "gs_set_walk": (I, [P, I]),
        "gs_set_plain_pairs": (I, [P, I, I]),
        "gs_set_fused_run": (I, [P, I]),
        "gs_set_diag_runs": (I, [P, I, I]),
        "gs_set_fold_reduce": (I, [P, I]),
        "gs_race_check": (I, [P, I, I, ctypes.POINTER(ctypes.c_int64)]),
        "gs_step_run": (I, [P, I, ip]),
        "gs_copy_probe": (I, [I, I64, dp]),
        "gs_batch_create": (I, [ctypes.POINTER(P), ip, I, I, I, pp]),
        "gs_batch_set_matrices": (I, [P, I, dp]),
        "gs_batch_grad": (I, [P, ctypes.POINTER(i8p), ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp), I, D,
                              ip]),
        "gs_batch_grad_ms": (D, [P, dp]),
        "gs_batch_destroy": (None, [P]),
        "gs_batch_score": (I, [P, ctypes.POINTER(i8p), ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp)]),
        "gs_batch_use_graph": (I, [P, I]),
        "gs_batch_set_tile": (I, [P, I, I, I]),
        "gs_batch_set_xcd_items": (I, [P, I]),
        "gs_batch_set_unroll": (I, [P, I]),
        "gs_tt_plan_load": (I, [ctypes.c_char_p, I, pp]),
        "gs_tt_plan_info": (I, [P, ip, ip, ip]),
        "gs_tt_set_grid_log2": (I, [P, I]),
        "gs_tt_plan_destroy": (None, [P]),
        "gs_tt_engine_create": (I, [P, I, I, pp]),
        "gs_tt_engine_destroy": (None, [P]),
        "gs_tt_prob": (I, [P, i8p, I64, dp]),
        "gs_tt_chain": (I, [P, i8p, I64, dp, dp, dp]),
        "gs_tt_sample": (I, [P, U64, I64, I64, i8p, dp]),
        "gs_tt_hybrid": (I, [P, i8p, I64, U64, I64, i8p, dp, dp]),
        "gs_tt_prefix": (I, [P, i8p, I, dp, I64, ip, dp, dp]),
        "gs_tt_resume": (I, [P, i8p, I, I, ctypes.POINTER(dp), dp, U64, I64, i8p, dp, dp, dp]),
        "gs_tt_dp": (I, [P, I64, ip, dp, dp, dp]),
        "gs_tt_last_error": (ctypes.c_char_p, []),
        "gs_batch_last_ms": (D, [P, dp]),
        "gs_batch_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc:
        raise GsError(rc, library().gs_last_error().decode())


def _i8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))


def _f8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def copy_probe(nbytes, device=0):
    """Duration (ms) of one plain device copy of nbytes (reads and writes nbytes): counter calibration."""
    ms = ctypes.c_double()
    _check(library().gs_copy_probe(device, nbytes, ctypes.byref(ms)))
    return ms.value


class Plan:
    """An in-place plan loaded by the engine's parser (same format and error codes as gsamp2.cpp:126-139)."""

    def __init__(self, path=None, text=None):
        lib = library()
        self._h = ctypes.c_void_p()
        if (path is None) == (text is None):
            raise ValueError("give exactly one of path, text")
        if path is not None:
            _check(lib.gs_plan_load(os.fsencode(path), ctypes.byref(self._h)))
            with open(path, "rb") as f:
                self.digest = hashlib.sha1(f.read()).hexdigest()
        else:
            _check(lib.gs_plan_parse(text.encode(), ctypes.byref(self._h)))
            self.digest = hashlib.sha1(text.encode()).hexdigest()
        n, ns, R = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.gs_plan_info(self._h, ctypes.byref(n), ctypes.byref(ns), ctypes.byref(R))
        self.n, self.ns, self.R = n.value, ns.value, R.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gs_plan_destroy(self._h)
            self._h = None


def prob_noalloc(plan, bits, storage=C64, device=0):
    """log2 Pr of forced strings like ``Engine.prob``, with no device allocation: the library's fixed 64 MiB device
    buffer holds the arena (R <= 22 in complex64, 21 in complex128).  Concurrent calls on one device race."""
    p = plan if isinstance(plan, Plan) else Plan(path=plan)
    b = np.ascontiguousarray(np.atleast_2d(np.asarray(bits, dtype=np.int8)))
    if b.shape[1] != p.ns:
        raise ValueError(f"strings must have {p.ns} steps, got {b.shape[1]}")
    out = np.empty(b.shape[0])
    _check(library().gs_prob_noalloc(p._h, storage, device, _i8(b), b.shape[0], _f8(out)))
    return out


class Engine:
    """A device arena of 2^R entries (``storage`` C64 or C128) bound to one plan on one HIP device."""

    def __init__(self, plan, storage=C64, device=0):
        lib = library()
        self.plan = plan if isinstance(plan, Plan) else Plan(path=plan)
        self.ns = self.plan.ns
        self.storage = storage
        self._h = ctypes.c_void_p()
        _check(lib.gs_engine_create(self.plan._h, storage, device, ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gs_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _bits(self, bits, allow_q=False):
        b = np.ascontiguousarray(np.atleast_2d(np.asarray(bits, dtype=np.int8)))
        if b.shape[1] != self.ns:
            raise ValueError(f"strings must have {self.ns} steps, got {b.shape[1]}")
        return b

    @property
    def last_device_ms(self):
        return library().gs_last_device_ms(self._h)

    def prob(self, bits):
        """log2 Pr of forced strings (``gsamp2 prob``)."""
        b = self._bits(bits)
        out = np.empty(b.shape[0])
        _check(library().gs_prob(self._h, _i8(b), b.shape[0], _f8(out)))
        return out

    def chain(self, bits):
        """``(log2 Pr, conditionals [nstr, ns], Z [nstr] complex)`` of forced strings (``gsamp2 chain``)."""
        b = self._bits(bits)
        lp = np.empty(b.shape[0])
        cond = np.empty((b.shape[0], self.ns))
        amp = np.empty((b.shape[0], 2))
        _check(library().gs_chain(self._h, _i8(b), b.shape[0], _f8(lp), _f8(cond), _f8(amp)))
        return lp, cond, amp[:, 0] + 1j * amp[:, 1]

    def sample(self, count, seed, first=0, seeding=SEED_PLAIN):
        """``(bits [count, ns], log2 Pr [count])`` of samples first..first+count-1 (``gsamp2 sample``/``sample2``)."""
        bits = np.empty((count, self.ns), np.int8)
        lp = np.empty(count)
        _check(library().gs_sample(self._h, seed, first, count, seeding, _i8(bits), _f8(lp)))
        return bits, lp

    def hybrid(self, bits, seed, first_line=0):
        """``(bits, log2 Pr, conditionals)``; -1 entries are sampled (``gsamp2 hybrid``)."""
        b = self._bits(bits, allow_q=True)
        ob = np.empty_like(b)
        lp = np.empty(b.shape[0])
        cond = np.empty((b.shape[0], self.ns))
        _check(library().gs_hybrid(self._h, _i8(b), b.shape[0], seed, first_line, _i8(ob), _f8(lp), _f8(cond)))
        return ob, lp, cond

    def prefix(self, bits, K):
        """Live table after the forced steps 0..K-1 of one string: ``(table [2^nk * ntq] complex, nk, ntq,
        logscale, log2 Pr(prefix))`` in the dump's logical order (``gsamp2 prefix``)."""
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8).reshape(-1))
        if b.size != self.ns:
            raise ValueError("one string of ns steps")
        nk, ntq = ctypes.c_int(), ctypes.c_int()
        lib = library()
        _check(lib.gs_prefix(self._h, _i8(b), K, None, 0, ctypes.byref(nk), ctypes.byref(ntq), None, None))
        n = (1 << nk.value) * ntq.value
        tab = np.empty(2 * n)
        ls, lp = ctypes.c_double(), ctypes.c_double()
        _check(lib.gs_prefix(self._h, _i8(b), K, _f8(tab), n, ctypes.byref(nk), ctypes.byref(ntq), ctypes.byref(ls),
                             ctypes.byref(lp)))
        return tab[0::2] + 1j * tab[1::2], nk.value, ntq.value, ls.value, lp.value

    def resume(self, bits, K, tables, logscales, seed, line=0, cshift=None, tqshift=None):
        """Sum the slice tables (logical order, with their logscales; optional twist), then run steps K..ns-1 with
        -1 entries sampled (``gsamp2 resume``).  Returns ``(bits, log2 Pr(full), log2 Pr(prefix), cond)``."""
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8).reshape(-1))
        if b.size != self.ns:
            raise ValueError("one string of ns steps")
        if not 0 <= K <= self.ns:
            raise ValueError("K out of range")
        nk, ntq = ctypes.c_int(), ctypes.c_int()        # the step-K table's shape (no steps run)
        _check(library().gs_prefix(self._h, _i8(np.zeros(self.ns, np.int8)), K, None, 0, ctypes.byref(nk),
                                   ctypes.byref(ntq), None, None))
        n = (1 << nk.value) * ntq.value
        tables = [np.asarray(t).reshape(-1) for t in tables]
        if not tables or any(t.size != n for t in tables):
            raise ValueError(f"every slice table must hold 2^nk * ntq = {n} entries")
        if len(logscales) != len(tables):
            raise ValueError("one logscale per slice table")
        if (cshift is None) != (tqshift is None) or (cshift is not None and (len(cshift) != len(tables) or
                                                                             len(tqshift) != len(tables))):
            raise ValueError("cshift and tqshift: both or neither, one entry per slice table")
        tabs = [np.ascontiguousarray(np.stack([np.real(t), np.imag(t)], axis=1).reshape(-1)) for t in tables]
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(tabs))(*[_f8(t) for t in tabs])
        ls = np.ascontiguousarray(np.asarray(logscales, dtype=np.float64))
        cs = ts = None
        if cshift is not None:
            cs_a = np.ascontiguousarray(np.asarray(cshift, dtype=np.uint64))
            ts_a = np.ascontiguousarray(np.asarray(tqshift, dtype=np.uint64))
            cs = cs_a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
            ts = ts_a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
        ob = np.empty(self.ns, np.int8)
        lpf, lpp = ctypes.c_double(), ctypes.c_double()
        cond = np.zeros(self.ns)
        _check(library().gs_resume(self._h, _i8(b), K, len(tabs), ptrs, _f8(ls), cs, ts, seed, line, _i8(ob),
                                   ctypes.byref(lpf), ctypes.byref(lpp), _f8(cond)))
        return ob, lpf.value, lpp.value, cond

    # profiling (engine extension, not a reference mode)
    def profile(self, on=True):
        _check(library().gs_profile_enable(self._h, 1 if on else 0))

    SWEEPS = ("gram", "direct", "write", "write+gram", "write+direct", "run sums", "run product")

    def profile_read(self):
        """``(ms [ns, 7], launches [ns, 7])`` per step and sweep kind (``SWEEPS``) since profiling was enabled (a run
        of diagonal steps counts at its first step)."""
        nk = len(self.SWEEPS)
        ms = np.zeros(self.ns * nk)
        n = np.zeros(self.ns * nk, np.int64)
        _check(library().gs_profile_read(self._h, _f8(ms), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return ms.reshape(self.ns, nk), n.reshape(self.ns, nk)

    def geometry(self, k):
        """Launch geometry of step k: dict(lr, grid, ntq, route, fusable, fgrid, route2, lg_iter, flr, flg_iter, vg,
        fvg)."""
        out = np.zeros(12, np.int32)
        _check(library().gs_step_geometry(self._h, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return dict(zip(("lr", "grid", "ntq", "route", "fusable", "fgrid", "route2", "lg_iter", "flr", "flg_iter", "vg",
                         "fvg"), (int(x) for x in out)))

    def columns(self, k, fused=False):
        """Coset-coordinate columns of step k as the plain or fused sweep uses them (None: no such layout)."""
        out = np.zeros(40, np.uint64)
        nk = library().gs_step_columns(self._h, k, 1 if fused else 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), 40)
        return None if nk < 0 else [int(x) for x in out[:nk]]

    def set_diag_runs(self, max_steps, min_nk=20):
        """Runs of at most max_steps (2..8; 0: off) consecutive diagonal steps of at least 2^min_nk cosets take one
        pattern-sum pass and one product pass (default 8, 20)."""
        _check(library().gs_set_diag_runs(self._h, int(max_steps), int(min_nk)))

    def run_of(self, k):
        """``(first step, steps, log2 blocks, log2 tiles, paired)`` of the run of diagonal steps holding step k, or
        None; paired: the run's product pass can also sum the next (two-slot) step's Gram matrix."""
        out = np.zeros(5, np.int32)
        _check(library().gs_step_run(self._h, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return None if out[0] < 0 else tuple(int(x) for x in out)

    RACE_LAYOUTS = ("plain", "fused", "logical", "run")

    def race_check(self, k, layout="plain"):
        """Device race check of step k's layout (overwrites the arena): dict(read_twice, written_twice,
        clobbered, checked); for a run of diagonal steps written_twice counts slots visited twice and clobbered the
        slots outside the step's table."""
        out = np.zeros(4, np.int64)
        _check(library().gs_race_check(self._h, k, self.RACE_LAYOUTS.index(layout),
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(zip(("read_twice", "written_twice", "clobbered", "checked"), (int(x) for x in out)))

    def race_check_all(self):
        """Race check of every layout the kernels walk (plain, fused where a step folds the next norms, runs of
        diagonal steps); returns the summed counts, and raises AssertionError on any conflict."""
        tot = dict(read_twice=0, written_twice=0, clobbered=0, checked=0)
        for k in range(self.ns):
            lays = ["plain"] + (["fused"] if self.geometry(k)["fusable"] else [])
            run = self.run_of(k)
            if run is not None and run[0] == k:
                lays.append("run")
            for lay in lays:
                r = self.race_check(k, lay)
                for key in tot:
                    tot[key] += r[key]
                if r["read_twice"] or r["written_twice"] or r["clobbered"]:
                    raise AssertionError(f"step {k} layout {lay}: {r}")
        return tot

    def set_fold_reduce(self, on):
        """Each sweep reduces the partial sums the previous sweep left (True), only the sweeps of steps with at most
        2^on live slots do (an int >= 2), or a one-block kernel does in between (False, default); bitwise the same
        results."""
        _check(library().gs_set_fold_reduce(self._h, int(on) if not isinstance(on, bool) else (1 if on else 0)))

    def set_fusion(self, on):
        """Fold each sampled step's norm pass into the previous write sweep where possible (default on)."""
        _check(library().gs_set_fusion(self._h, 1 if on else 0))

    def set_fused4(self, on):
        """Fold a four-slot next step's norms into the previous write too (default on)."""
        _check(library().gs_set_fused4(self._h, 1 if on else 0))

    def set_fused_run(self, lr):
        """log2 run length of the fused layout (complex64: 1 or 2; -1: default)."""
        _check(library().gs_set_fused_run(self._h, int(lr)))

    def set_walk(self, mode):
        """Tile order of the sweeps: 0 forward, 1 reverse on odd steps, 2 reverse on every step."""
        _check(library().gs_set_walk(self._h, int(mode)))

    def set_plain_pairs(self, ntq_max, far=0):
        """Plain sweeps of steps with at most ntq_max slots per coset (0..2; default 0, off) take two runs per thread,
        the second at the top coordinate (far=1) or at the next unit bit (far=0)."""
        _check(library().gs_set_plain_pairs(self._h, int(ntq_max), int(far)))

    def set_xcd_blocks(self, on):
        """Each XCD takes a contiguous eighth of every sweep's block coordinates (True, default) or blocks go
        round-robin."""
        _check(library().gs_set_xcd_blocks(self._h, 1 if on else 0))

    def set_small_steps(self, nk):
        """Run consecutive steps of at most 2^nk cosets in one single-block launch (-1: never)."""
        _check(library().gs_set_small_steps(self._h, int(nk)))

    def set_grid_log2_one_slot(self, lg):
        """Cap the plain sweeps of one-slot steps alone at 2^lg blocks (1..14; default 12, or 14 from 2^30 cosets
        on; 0 restores those defaults)."""
        _check(library().gs_set_grid_log2_one_slot(self._h, int(lg)))

    def set_grid_log2(self, lg):
        """Cap every sweep at 2^lg blocks (1..14; default 9, one-slot plain sweeps 12, or 14 from 2^30 cosets on);
        tiles beyond the cap loop inside each block.  Sets the one-slot caps to lg as well; 0 restores every
        default."""
        _check(library().gs_set_grid_log2(self._h, int(lg)))


class BatchEngine:
    """Batched chain-mode scoring of many strings of many plans, one launch per step (``gs_batch_*``).

    ``plans``: list of Plan objects or paths; ``lanes``: strings per plan per call (multiples of 64)."""

    def __init__(self, plans, lanes, storage=C64, device=0):
        lib = library()
        self.plans = [p if isinstance(p, Plan) else Plan(path=p) for p in plans]
        self.lanes = [int(x) for x in lanes]
        self.storage = storage
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h for p in self.plans])
        ln = (ctypes.c_int * len(self.lanes))(*self.lanes)
        self._h = ctypes.c_void_p()
        rc = lib.gs_batch_create(arr, ln, len(self.plans), storage, device, ctypes.byref(self._h))
        if rc:
            raise GsError(rc, lib.gs_batch_last_error().decode())
        self.maxns = max(p.ns for p in self.plans)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc:
            raise GsError(rc, library().gs_batch_last_error().decode())

    def use_graph(self, on=True):
        self._check(library().gs_batch_use_graph(self._h, 1 if on else 0))

    def set_tile(self, i, lg, k=-1):
        self._check(library().gs_batch_set_tile(self._h, i, k, lg))

    def set_xcd_items(self, on):
        """Work-list order: 0 round-robin over the XCDs, 1 (default) each XCD takes consecutive items (lane group by
        lane group), 2 each XCD takes consecutive coset ranges with all their lane groups."""
        self._check(library().gs_batch_set_xcd_items(self._h, int(on)))

    def set_unroll(self, lu):
        """log2 cosets each lane gathers at once (0..3)."""
        self._check(library().gs_batch_set_unroll(self._h, int(lu)))

    def score(self, bits, want_cond=True, want_amp=True):
        """bits: list (per plan) of int8 [lanes_i, ns_i].  Returns lists of log2p [lanes_i], cond [lanes_i, ns_i],
        amp [lanes_i] complex (cond / amp entries are None when not requested)."""
        n = len(self.plans)
        if len(bits) != n:
            raise ValueError(f"one bit array per plan ({n}), got {len(bits)}")
        bs = [np.ascontiguousarray(np.asarray(b, dtype=np.int8)) for b in bits]
        for b, p, ln in zip(bs, self.plans, self.lanes):
            if b.shape != (ln, p.ns):
                raise ValueError(f"bits must be [{ln}, {p.ns}], got {b.shape}")
        lp = [np.empty(ln) for ln in self.lanes]
        cd = [np.empty((ln, p.ns)) for ln, p in zip(self.lanes, self.plans)]
        am = [np.empty((ln, 2)) for ln in self.lanes]
        I8P = ctypes.POINTER(ctypes.c_int8)
        DP = ctypes.POINTER(ctypes.c_double)
        bp = (I8P * n)(*[_i8(b) for b in bs])
        lpp = (DP * n)(*[_f8(x) for x in lp])
        cdp = (DP * n)(*[_f8(x) for x in cd]) if want_cond else None
        amp = (DP * n)(*[_f8(x) for x in am]) if want_amp else None
        self._check(library().gs_batch_score(self._h, bp, lpp, cdp, amp))
        return (lp, cd if want_cond else [None] * n,
                [a[:, 0] + 1j * a[:, 1] for a in am] if want_amp else [None] * n)

    def last_ms(self):
        """``(device ms of the last score, end of each step in ms after its start or -1)``."""
        st = np.zeros(self.maxns)
        ms = library().gs_batch_last_ms(self._h, _f8(st))
        return ms, st

    def set_matrices(self, i, M):
        """Instance i's measurement matrices: complex array [ns, 2, 2] (M[k, b, x])."""
        M = np.asarray(M, dtype=complex)
        if M.shape != (self.plans[i].ns, 2, 2):
            raise ValueError("M must be [ns, 2, 2]")
        flat = np.ascontiguousarray(np.stack([M.real, M.imag], axis=-1).reshape(-1))
        self._check(library().gs_batch_set_matrices(self._h, i, _f8(flat)))

    def grad(self, bits, seg=0, mem_budget=0):
        """Adjoint: per plan ``(log2p [B], Z [B] complex, G [B, ns, 2] complex)`` with G[b, k, x] = dZ_b/dM_k[b_k, x];
        checkpoint segments of ``seg`` steps (0: the segment that recomputes least within ``mem_budget`` bytes of
        copies; budget 0: 70 % of the free device memory).  Returns (results, seg)."""
        n = len(self.plans)
        if len(bits) != n:
            raise ValueError(f"one bit array per plan ({n}), got {len(bits)}")
        bs = [np.ascontiguousarray(np.asarray(b, dtype=np.int8)) for b in bits]
        for b, p, ln in zip(bs, self.plans, self.lanes):
            if b.shape != (ln, p.ns):
                raise ValueError(f"bits must be [{ln}, {p.ns}], got {b.shape}")
        lp = [np.empty(ln) for ln in self.lanes]
        am = [np.empty((ln, 2)) for ln in self.lanes]
        G = [np.empty((ln, p.ns, 2, 2)) for ln, p in zip(self.lanes, self.plans)]
        I8P = ctypes.POINTER(ctypes.c_int8)
        DP = ctypes.POINTER(ctypes.c_double)
        used = ctypes.c_int()
        self._check(library().gs_batch_grad(self._h, (I8P * n)(*[_i8(b) for b in bs]), (DP * n)(*[_f8(x) for x in lp]),
                                            (DP * n)(*[_f8(x) for x in am]), (DP * n)(*[_f8(x) for x in G]), seg,
                                            mem_budget, ctypes.byref(used)))
        res = [(lp[i], am[i][:, 0] + 1j * am[i][:, 1], G[i][..., 0] + 1j * G[i][..., 1]) for i in range(n)]
        return res, used.value

    def grad_ms(self):
        """``(total, forward)`` device ms of the last grad call."""
        f = ctypes.c_double()
        t = library().gs_batch_grad_ms(self._h, ctypes.byref(f))
        return t, f.value


class TwoTableEngine:
    """The two-table (gather-form) engine: the reference's gsamp.cpp sampler (``with_m=True`` plans) and dp.cpp amplitude
    engine (``with_m=False``); two tables of 2^maxr entries on one device."""

    def __init__(self, path, with_m=True, storage=C64, device=0):
        lib = library()
        self._plan = ctypes.c_void_p()
        self._h = None
        rc = lib.gs_tt_plan_load(os.fsencode(path), 1 if with_m else 0, ctypes.byref(self._plan))
        if rc:
            raise GsError(rc, lib.gs_tt_last_error().decode())
        n, ns, maxr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.gs_tt_plan_info(self._plan, ctypes.byref(n), ctypes.byref(ns), ctypes.byref(maxr))
        self.n, self.ns, self.maxr = n.value, ns.value, maxr.value
        h = ctypes.c_void_p()
        self._check(lib.gs_tt_engine_create(self._plan, storage, device, ctypes.byref(h)))
        self._h = h

    def _check(self, rc):
        if rc:
            raise GsError(rc, library().gs_tt_last_error().decode())

    def set_grid_log2(self, lg):
        """Cap every gather sweep at 2^lg blocks (1..14; 0: the default)."""
        self._check(library().gs_tt_set_grid_log2(self._h, int(lg)))

    def close(self):
        if _lib is not None:
            if getattr(self, "_h", None):
                _lib.gs_tt_engine_destroy(self._h)
                self._h = None
            if getattr(self, "_plan", None):
                _lib.gs_tt_plan_destroy(self._plan)
                self._plan = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _bits(self, bits, allow_q=False):
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8))
        if b.ndim == 1:
            b = b[None, :]
        if b.shape[1] != self.ns:
            raise ValueError(f"strings must have {self.ns} steps")
        return b

    def prob(self, bits):
        b = self._bits(bits)
        out = np.empty(b.shape[0])
        self._check(library().gs_tt_prob(self._h, _i8(b), b.shape[0], _f8(out)))
        return out

    def chain(self, bits):
        """``(log2 Pr, conditionals [nstr, ns], Z [nstr] complex)`` (``gsamp chain``)."""
        b = self._bits(bits)
        lp = np.empty(b.shape[0])
        cond = np.empty((b.shape[0], self.ns))
        amp = np.empty((b.shape[0], 2))
        self._check(library().gs_tt_chain(self._h, _i8(b), b.shape[0], _f8(lp), _f8(cond), _f8(amp)))
        return lp, cond, amp[:, 0] + 1j * amp[:, 1]

    def sample(self, count, seed, first=0):
        bits = np.empty((count, self.ns), np.int8)
        lp = np.empty(count)
        self._check(library().gs_tt_sample(self._h, ctypes.c_uint64(seed), first, count, _i8(bits), _f8(lp)))
        return bits, lp

    def hybrid(self, bits, seed, first_line=0):
        b = self._bits(bits, allow_q=True)
        ob = np.empty_like(b)
        lp = np.empty(b.shape[0])
        cond = np.empty((b.shape[0], self.ns))
        self._check(library().gs_tt_hybrid(self._h, _i8(b), b.shape[0], ctypes.c_uint64(seed), first_line, _i8(ob),
                                           _f8(lp), _f8(cond)))
        return ob, lp, cond

    def prefix(self, bits, K):
        """``(table [2^r] complex, r, logscale, log2 Pr(prefix))`` after the forced steps 0..K-1 (``gsamp prefix``)."""
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8).reshape(-1))
        r = ctypes.c_int()
        lib = library()
        self._check(lib.gs_tt_prefix(self._h, _i8(b), K, None, 0, ctypes.byref(r), None, None))
        n = 1 << r.value
        tab = np.empty(2 * n)
        ls, lp = ctypes.c_double(), ctypes.c_double()
        self._check(lib.gs_tt_prefix(self._h, _i8(b), K, _f8(tab), n, ctypes.byref(r), ctypes.byref(ls),
                                     ctypes.byref(lp)))
        return tab[0::2] + 1j * tab[1::2], r.value, ls.value, lp.value

    def resume(self, bits, K, tables, logscales, seed, line=0):
        """``(bits, log2 Pr(full), log2 Pr(prefix), cond)`` (``gsamp resume``; -1 entries are sampled)."""
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8).reshape(-1))
        tabs = [np.ascontiguousarray(np.stack([np.real(t), np.imag(t)], axis=1).reshape(-1)) for t in tables]
        ptrs = (ctypes.POINTER(ctypes.c_double) * len(tabs))(*[_f8(t) for t in tabs])
        ls = np.ascontiguousarray(np.asarray(logscales, dtype=np.float64))
        ob = np.empty(self.ns, np.int8)
        lpf, lpp = ctypes.c_double(), ctypes.c_double()
        cond = np.zeros(self.ns)
        self._check(library().gs_tt_resume(self._h, _i8(b), K, len(tabs), ptrs, _f8(ls), ctypes.c_uint64(seed), line,
                                           _i8(ob), ctypes.byref(lpf), ctypes.byref(lpp), _f8(cond)))
        return ob, lpf.value, lpp.value, cond

    def dp(self, jobs):
        """jobs: [[(pre complex, [(w0, w1)] per vertex)] per job] -> complex [njobs] (``dp``)."""
        nsl = np.array([len(j) for j in jobs], np.int32)
        pre = np.array([[p.real, p.imag] for j in jobs for p, _ in j], np.float64).reshape(-1)
        w = np.array([[[a.real, a.imag, c.real, c.imag] for a, c in ws] for j in jobs for _, ws in j],
                     np.float64).reshape(-1)
        out = np.empty(2 * len(jobs))
        self._check(library().gs_tt_dp(self._h, len(jobs), nsl.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       _f8(np.ascontiguousarray(pre)), _f8(np.ascontiguousarray(w)), _f8(out)))
        return out[0::2] + 1j * out[1::2]
