"""Python host API of the MI355X propagator-completion engine (ctypes over libpcx.so).

Mirrors the reference's PIR surface (lattice-land/lala-pc include/lala/pir.hpp):

    PIR.interpret_tell / PIR.tell   -> pcx_tell       (pir.hpp:316-352)
    PIR.num_deductions              -> pcx_num_deductions (pir.hpp:382-384)
    PIR.fixpoint                    -> pcx_fixpoint   (GaussSeidelIteration over
                                                        PIR::deduce, pir_test.cpp:60-62)
    PIR.ask_all / PIR.is_extractable-> pcx_ask_all    (pir.hpp:368-370, 873-884)
    Batch                           -> shared table, one store per search node
                                        (pir.hpp:182-205 shared-copy clones)

Every computation runs on the GPU through the compiled library; there is no
host fallback.  Loading the library needs no GPU (the generators are host
code); creating an engine without a HIP device raises PcxError.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcx.so")

INF = 2**31 - 1
MINF = -(2**31)

ADD, MUL, TDIV, FDIV, CDIV, EDIV, MIN, MAX, EQ, LEQ = range(10)
OPS = {"ADD": ADD, "MUL": MUL, "TDIV": TDIV, "FDIV": FDIV, "CDIV": CDIV, "EDIV": EDIV,
       "MIN": MIN, "MAX": MAX, "EQ": EQ, "LEQ": LEQ}

MODE_PLAIN, MODE_BATCHED, MODE_PERSISTENT = 0, 1, 2

DOMAIN_INTERVAL, DOMAIN_BITSET = 0, 1
BLEQ, BADD, BNEQ, BEQR = range(4)
BITS_OPS = {"LEQ": BLEQ, "ADD": BADD, "NEQ": BNEQ, "EQR": BEQR}

OK, EINVAL, ENOMEM, EDEVICE, EBUSY, ENOTSUP = range(6)


class PcxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"pcx status {status}: {msg}")
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("props_per_wave", ctypes.c_int32),
                ("lds_mode", ctypes.c_int32), ("sweeps_per_check", ctypes.c_int32),
                ("max_sweeps", ctypes.c_int64), ("timing", ctypes.c_int32),
                ("domain", ctypes.c_int32), ("compact_tables", ctypes.c_int32),
                ("stream_tables", ctypes.c_int32), ("event_driven", ctypes.c_int32),
                ("lockstep", ctypes.c_int32), ("mark_queue", ctypes.c_int32),
                ("base_seed", ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [("has_changed", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("sweeps", ctypes.c_int64), ("evals", ctypes.c_int64)]


RESULT_DTYPE = np.dtype([("has_changed", np.int32), ("failed", np.int32), ("sweeps", np.int64),
                         ("evals", np.int64)])


class Stats(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("kernel_launches", ctypes.c_int64), ("evals", ctypes.c_int64),
                ("tightened", ctypes.c_int64), ("sweeps_max", ctypes.c_int64),
                ("full_ms", ctypes.c_double), ("full_launches", ctypes.c_int64),
                ("full_evals", ctypes.c_int64), ("full_tightened", ctypes.c_int64),
                ("lockstep", ctypes.c_int64), ("load_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GenParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("nvars", ctypes.c_int64), ("nprops", ctypes.c_int64),
                ("kinds", ctypes.c_uint32), ("planted", ctypes.c_int32),
                ("lb_max", ctypes.c_int32), ("width_min", ctypes.c_int32),
                ("width_max", ctypes.c_int32), ("value_range", ctypes.c_int32),
                ("bool_frac", ctypes.c_double), ("const_frac", ctypes.c_double),
                ("chain_frac", ctypes.c_double), ("chain_len", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 3)]


class Tuning(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("props_per_wave", ctypes.c_int32),
                ("lds_mode", ctypes.c_int32), ("compact_tables", ctypes.c_int32),
                ("stream_tables", ctypes.c_int32), ("event_driven", ctypes.c_int32),
                ("lockstep", ctypes.c_int32), ("ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_slots", ctypes.c_int32),
                ("arena_bytes", ctypes.c_int64), ("sweeps_per_tick", ctypes.c_int32),
                ("props_per_wave", ctypes.c_int32), ("stage_bytes", ctypes.c_int64),
                ("event_driven", ctypes.c_int32), ("burst_sweeps", ctypes.c_int32),
                ("burst_after", ctypes.c_int32), ("reserved", ctypes.c_int32 * 1)]


class ServerStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_int64), ("sweeps", ctypes.c_int64), ("admitted", ctypes.c_int64),
                ("completed", ctypes.c_int64), ("waits_slot", ctypes.c_int64),
                ("waits_arena", ctypes.c_int64), ("arena_peak_bytes", ctypes.c_int64),
                ("max_running", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("waits_stage_admit", ctypes.c_int64), ("waits_stage_harvest", ctypes.c_int64),
                ("waits_stage_side", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Served(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int64), ("result", Result), ("latency_ms", ctypes.c_double)]


_P32 = ctypes.POINTER(ctypes.c_int32)
_P64 = ctypes.POINTER(ctypes.c_int64)
_PU8 = ctypes.POINTER(ctypes.c_uint8)
_PU64 = ctypes.POINTER(ctypes.c_uint64)
_VP = ctypes.c_void_p

# (name, restype, argtypes) of every exported C symbol (include/pcx.h, include/pcx_gen.h)
SIGNATURES = [
    ("pcx_version", ctypes.c_char_p, []),
    ("pcx_last_error", ctypes.c_char_p, []),
    ("pcx_config_default", None, [ctypes.POINTER(Config)]),
    ("pcx_device_count", ctypes.c_int32, []),
    ("pcx_create", ctypes.c_int32, [ctypes.POINTER(Config), ctypes.POINTER(_VP)]),
    ("pcx_destroy", None, [_VP]),
    ("pcx_tell", ctypes.c_int32, [_VP, _P32, _P32, _P32, _P32, ctypes.c_int64, _P32, _P32, ctypes.c_int64]),
    ("pcx_num_deductions", ctypes.c_int64, [_VP]),
    ("pcx_get_config", ctypes.c_int32, [_VP, ctypes.POINTER(Config)]),
    ("pcx_set_config", ctypes.c_int32, [_VP, ctypes.POINTER(Config)]),
    ("pcx_get_table", ctypes.c_int32, [_VP, _P32, _P32, _P32, _P32]),
    ("pcx_snapshot", ctypes.c_int64, [_VP]),
    ("pcx_restore", ctypes.c_int32, [_VP, ctypes.c_int64]),
    ("pcx_fixpoint", ctypes.c_int32, [_VP, _P32, _P32, ctypes.c_int64, ctypes.POINTER(Result)]),
    ("pcx_fixpoint_noalloc", ctypes.c_int32, [_VP, _VP, _VP, ctypes.c_int64, ctypes.POINTER(Result)]),
    ("pcx_noalloc_max_vars", ctypes.c_int64, []),
    ("pcx_ask_all", ctypes.c_int32, [_VP, _P32, _P32, ctypes.c_int64, _PU8, _P32]),
    ("pcx_batch_create", ctypes.c_int32, [_VP, ctypes.c_int64, ctypes.c_int64, _P64, _P32, ctypes.POINTER(_VP)]),
    ("pcx_batch_destroy", None, [_VP]),
    ("pcx_batch_load", ctypes.c_int32, [_VP, _P32, _P32]),
    ("pcx_batch_load_device", ctypes.c_int32, [_VP, _VP, _VP]),
    ("pcx_batch_reset", ctypes.c_int32, [_VP]),
    ("pcx_batch_read", ctypes.c_int32, [_VP, _P32, _P32]),
    ("pcx_batch_run", ctypes.c_int32, [_VP, ctypes.c_int32, ctypes.POINTER(Result), ctypes.POINTER(Stats)]),
    ("pcx_batch_branch", ctypes.c_int32, [_VP, _VP, _P32, ctypes.c_int32]),
    ("pcx_batch_ask", ctypes.c_int32, [_VP, _P32, _P64, _P32, ctypes.c_int64]),
    ("pcx_batch_ask_formulas", ctypes.c_int32, [_VP, _P32, ctypes.c_int64, _P64, ctypes.c_int64, _P32]),
    ("pcx_batch_autotune", ctypes.c_int32, [_VP, ctypes.c_int32, ctypes.POINTER(Tuning),
                                            ctypes.POINTER(Tuning), ctypes.c_int32, _P32]),
    ("pcx_bits_tell", ctypes.c_int32, [_VP, _P32, _P32, _P32, _P32, ctypes.c_int64, _PU64, ctypes.c_int64]),
    ("pcx_bits_fixpoint", ctypes.c_int32, [_VP, _PU64, ctypes.c_int64, ctypes.POINTER(Result)]),
    ("pcx_bits_batch_load", ctypes.c_int32, [_VP, _PU64]),
    ("pcx_bits_batch_read", ctypes.c_int32, [_VP, _PU64]),
    ("pcx_server_config_default", None, [ctypes.POINTER(ServerConfig)]),
    ("pcx_server_create", ctypes.c_int32, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]),
    ("pcx_server_destroy", None, [_VP]),
    ("pcx_server_add_network", ctypes.c_int32, [_VP, _VP, _P32]),
    ("pcx_server_set_root", ctypes.c_int32, [_VP, ctypes.c_int32, _P32, _P32, ctypes.c_int64]),
    ("pcx_server_submit", ctypes.c_int32, [_VP, ctypes.c_int32, _P32, _P32, ctypes.c_int64,
                                           ctypes.c_int64]),
    ("pcx_server_step", ctypes.c_int32, [_VP, _P32]),
    ("pcx_server_get_stats", ctypes.c_int32, [_VP, ctypes.POINTER(ServerStats)]),
    ("pcx_server_poll", ctypes.c_int64, [_VP, ctypes.POINTER(Served), ctypes.c_int64]),
    ("pcx_server_replay", ctypes.c_int32, [_VP, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), _P32,
                                           _P32, _P32, _P64, ctypes.c_int32, ctypes.POINTER(Served),
                                           ctypes.POINTER(ctypes.c_double)]),
    ("pcxg_params_default", None, [ctypes.POINTER(GenParams)]),
    ("pcxg_network", ctypes.c_int32, [ctypes.POINTER(GenParams), _P32, _P32, _P32, _P32, _P32, _P32, _P32]),
    ("pcxg_splits", ctypes.c_int32, [ctypes.c_uint64, _P32, _P32, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int32, ctypes.c_int32, _P32, _P32]),
    ("pcxg_subsets", ctypes.c_int32, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_int64, _P64, _P32]),
    ("pcxg_bits_network", ctypes.c_int32, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                           _P32, _P32, _P32, _P32, _PU64, _P32]),
    ("pcxg_bits_splits", ctypes.c_int32, [ctypes.c_uint64, _PU64, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_int32, ctypes.c_int32, _PU64]),
    ("pcx_model_parse", ctypes.c_int32, [ctypes.c_char_p, ctypes.POINTER(_VP)]),
    ("pcx_model_add", ctypes.c_int32, [_VP, ctypes.c_char_p]),
    ("pcx_model_destroy", None, [_VP]),
    ("pcx_model_num_vars", ctypes.c_int64, [_VP]),
    ("pcx_model_num_props", ctypes.c_int64, [_VP]),
    ("pcx_model_table", ctypes.c_int32, [_VP, ctypes.c_int64, _P32, _P32, _P32, _P32]),
    ("pcx_model_store", ctypes.c_int32, [_VP, _P32, _P32]),
    ("pcx_model_var", ctypes.c_int32, [_VP, ctypes.c_char_p]),
    ("pcx_model_var_name", ctypes.c_char_p, [_VP, ctypes.c_int32]),
    ("pcx_model_num_formulas", ctypes.c_int64, [_VP]),
    ("pcx_model_num_formula_nodes", ctypes.c_int64, [_VP]),
    ("pcx_model_formulas", ctypes.c_int32, [_VP, _P32, _P64]),
]

_lib = None


def lib():
    """Load libpcx.so (built by __graft_entry__.build() / `make -C lala-pc_amd`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PcxError(EDEVICE, f"{LIB_PATH} is not built; run `make -C lala-pc_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a, dtype, ptype, size=None):
    """ctypes pointer to a C-contiguous array of `dtype` holding at least
    `size` elements.  The C side reads and writes exactly the sizes it is
    given, so a wrong dtype, layout or length is refused here (ValueError),
    not left to corrupt memory."""
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{np.dtype(dtype).name} C-contiguous numpy array expected")
    if size is not None and a.size < size:
        raise ValueError(f"array of {a.size} elements where {size} are needed")
    return a.ctypes.data_as(ptype)


def _p32(a, size=None):
    return _ptr(a, np.int32, _P32, size)


def _p64(a, size=None):
    return _ptr(a, np.int64, _P64, size)


def _pu64(a, size=None):
    return _ptr(a, np.uint64, _PU64, size)


def _store_pair(lb, ub):
    """Length of an lb/ub store pair, refusing a mismatch."""
    if lb is None or ub is None:
        if lb is not None or ub is not None:
            raise ValueError("lb and ub must both be given or both be None")
        return 0
    if lb.size != ub.size:
        raise ValueError(f"lb has {lb.size} entries and ub {ub.size}")
    return lb.size


def _check(st):
    if st != OK:
        raise PcxError(st, lib().pcx_last_error().decode(errors="replace"))


def device_count():
    return int(lib().pcx_device_count())


def default_config(**kw):
    c = Config()
    lib().pcx_config_default(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class PIR:
    """Propagator table on one GPU (the reference's PIR without its store).

    Stores are numpy int32 arrays lb/ub owned by the caller, as the
    reference mutates its sub-store in place.
    """

    def __init__(self, config=None, **kw):
        cfg = config if config is not None else default_config(**kw)
        h = _VP()
        _check(lib().pcx_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.config = cfg

    def close(self):
        if getattr(self, "_h", None):
            lib().pcx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tell(self, op, x, y, z, lb=None, ub=None):
        """Append propagators x = y <op> z; embeds reified outputs in [0,1] in lb/ub."""
        op, x, y, z = (np.ascontiguousarray(a, dtype=np.int32) for a in (op, x, y, z))
        n = len(op)
        if not (len(x) == len(y) == len(z) == n):
            raise ValueError("op, x, y, z must have the same length")
        nv = _store_pair(lb, ub)
        _check(lib().pcx_tell(self._h, _p32(op), _p32(x), _p32(y), _p32(z), n, _p32(lb, nv),
                              _p32(ub, nv), nv))

    interpret_tell = tell

    def num_deductions(self):
        return int(lib().pcx_num_deductions(self._h))

    def get_config(self):
        c = Config()
        _check(lib().pcx_get_config(self._h, ctypes.byref(c)))
        return c

    def set_config(self, **kw):
        """Change tuning fields (props_per_wave, lds_mode, compact_tables, ...)."""
        c = self.get_config()
        for k, v in kw.items():
            setattr(c, k, v)
        _check(lib().pcx_set_config(self._h, ctypes.byref(c)))
        self.config = c

    def snapshot(self):
        """The table size (pir.hpp:857-860); stores are the caller's."""
        return int(lib().pcx_snapshot(self._h))

    def restore(self, snap):
        """Pop the propagators told since `snap` off the end of the sorted
        table, as the reference does (pir.hpp:862-870)."""
        _check(lib().pcx_restore(self._h, int(snap)))

    def table(self):
        n = self.num_deductions()
        out = [np.zeros(n, np.int32) for _ in range(4)]
        _check(lib().pcx_get_table(self._h, *(_p32(a) for a in out)))
        return out

    def fixpoint(self, lb, ub):
        """Propagate lb/ub (int32, modified in place) to the fixpoint; returns Result."""
        r = Result()
        nv = _store_pair(lb, ub)
        _check(lib().pcx_fixpoint(self._h, _p32(lb, nv), _p32(ub, nv), nv, ctypes.byref(r)))
        return r

    def fixpoint_noalloc(self, d_lb, d_ub, nvars):
        """pcx_fixpoint_noalloc: the fixpoint of a store held in device memory
        (d_lb / d_ub: device addresses of nvars int32 bounds each, updated in
        place); the call allocates nothing.  Returns Result."""
        if not d_lb or not d_ub:
            raise ValueError("device addresses expected")
        r = Result()
        _check(lib().pcx_fixpoint_noalloc(self._h, _VP(int(d_lb)), _VP(int(d_ub)), int(nvars),
                                          ctypes.byref(r)))
        return r

    def ask_all(self, lb, ub):
        n = self.num_deductions()
        ent = np.zeros(max(n, 1), np.uint8)
        allf = ctypes.c_int32(0)
        nv = _store_pair(lb, ub)
        _check(lib().pcx_ask_all(self._h, _p32(lb, nv), _p32(ub, nv), nv,
                                 ent.ctypes.data_as(_PU8), ctypes.byref(allf)))
        return ent[:n].astype(bool), bool(allf.value)

    def is_extractable(self, lb, ub):
        """Not failed and every propagator entailed (pir.hpp:873-884)."""
        if (lb > ub).any():
            return False
        return self.ask_all(lb, ub)[1]


class BitsPIR(PIR):
    """Bitset-domain propagators (PC over 64-value bitset words).

    Stores are numpy uint64 arrays, one word per variable (bit v = value v).
    """

    def __init__(self, config=None, **kw):
        kw.setdefault("domain", DOMAIN_BITSET)
        super().__init__(config, **kw)

    def tell(self, op, x, y, z, words=None):
        op, x, y, z = (np.ascontiguousarray(a, dtype=np.int32) for a in (op, x, y, z))
        if not (len(x) == len(y) == len(z) == len(op)):
            raise ValueError("op, x, y, z must have the same length")
        nv = 0 if words is None else words.size
        _check(lib().pcx_bits_tell(self._h, _p32(op), _p32(x), _p32(y), _p32(z), len(op),
                                   _pu64(words), nv))

    interpret_tell = tell

    def fixpoint(self, words):
        r = Result()
        _check(lib().pcx_bits_fixpoint(self._h, _pu64(words), words.size, ctypes.byref(r)))
        return r


class Batch:
    """Many device-resident stores over one shared PIR table."""

    def __init__(self, pir, n_inst, nvars, inst_off=None, prop_idx=None):
        self.pir = pir
        self.n_inst = int(n_inst)
        self.nvars = int(nvars)
        n_idx = None
        if inst_off is not None:
            inst_off = np.ascontiguousarray(inst_off, dtype=np.int64)
            prop_idx = np.ascontiguousarray(prop_idx, dtype=np.int32)
            if inst_off.size != self.n_inst + 1:
                raise ValueError(f"inst_off needs n_inst + 1 = {self.n_inst + 1} entries")
            n_idx = int(inst_off[-1])
        h = _VP()
        _check(lib().pcx_batch_create(pir._h, self.n_inst, self.nvars, _p64(inst_off),
                                      _p32(prop_idx, n_idx), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().pcx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, lb, ub):
        n = self.n_inst * self.nvars
        if _store_pair(lb, ub) != n:
            raise ValueError(f"a batch load needs n_inst * nvars = {n} bounds in lb and in ub")
        _check(lib().pcx_batch_load(self._h, _p32(lb.reshape(-1), n), _p32(ub.reshape(-1), n)))

    def load_device(self, d_lb, d_ub):
        """Load the stores from device memory: d_lb / d_ub are device addresses
        (ints) of n_inst * nvars int32 bounds each, laid out as for load()."""
        if not d_lb or not d_ub:
            raise ValueError("device addresses expected")
        _check(lib().pcx_batch_load_device(self._h, _VP(int(d_lb)), _VP(int(d_ub))))

    def reset(self):
        _check(lib().pcx_batch_reset(self._h))

    def load_words(self, words):
        n = self.n_inst * self.nvars
        if words.size != n:
            raise ValueError(f"a batch load needs n_inst * nvars = {n} words")
        _check(lib().pcx_bits_batch_load(self._h, _pu64(words.reshape(-1), n)))

    def read_words(self):
        w = np.zeros(self.n_inst * self.nvars, np.uint64)
        _check(lib().pcx_bits_batch_read(self._h, _pu64(w)))
        return w

    def read(self):
        lb = np.zeros(self.n_inst * self.nvars, np.int32)
        ub = np.zeros(self.n_inst * self.nvars, np.int32)
        _check(lib().pcx_batch_read(self._h, _p32(lb), _p32(ub)))
        return lb, ub

    def run(self, mode=MODE_BATCHED):
        """Propagate every instance; returns (results, Stats).  results is a
        numpy record array (fields has_changed, failed, sweeps, evals; one
        record per instance, attribute access r.failed)."""
        res = np.empty(self.n_inst, RESULT_DTYPE)
        st = Stats()
        _check(lib().pcx_batch_run(self._h, mode, res.ctypes.data_as(ctypes.POINTER(Result)),
                                   ctypes.byref(st)))
        return res.view(np.recarray), st

    def branch(self, child, from_fixpoint=False):
        """Split every instance on its first unfixed variable into child
        instances 2k / 2k+1 (child: a Batch of 2 * n_inst instances over the
        same engine); returns the branching variables (-1: no split).
        from_fixpoint: the parents were run to their fixpoints, so the
        children's next run starts from the split variable's propagators."""
        var = np.zeros(max(self.n_inst, 1), np.int32)
        _check(lib().pcx_batch_branch(self._h, child._h, _p32(var), 1 if from_fixpoint else 0))
        return var[:self.n_inst]

    def ask_formulas(self, nodes, off):
        """PC's is_extractable of every store for formula trees (nodes an
        (n, 4) int32 array, off the n_formulas + 1 offsets, as
        Model.formulas() gives them): a bool array, one per instance."""
        nodes = np.ascontiguousarray(nodes, np.int32).reshape(-1, 4)
        off = np.ascontiguousarray(off, np.int64)
        out = np.zeros(self.n_inst, np.int32)
        _check(lib().pcx_batch_ask_formulas(self._h, nodes.ctypes.data_as(_P32), len(nodes), _p64(off),
                                            len(off) - 1, _p32(out)))
        return out.astype(bool)

    def ask(self, live=True):
        """Entailment on the current stores: (extractable bool[n_inst],
        live_off int64[n_inst+1], live_idx int32[...]) -- the propagators of
        each instance that are not entailed (None, None without `live`)."""
        ext = np.zeros(self.n_inst, np.int32)
        if not live:
            _check(lib().pcx_batch_ask(self._h, _p32(ext), None, None, 0))
            return ext.astype(bool), None, None
        off = np.zeros(self.n_inst + 1, np.int64)
        _check(lib().pcx_batch_ask(self._h, _p32(ext), _p64(off), None, 0))
        idx = np.zeros(max(int(off[-1]), 1), np.int32)
        _check(lib().pcx_batch_ask(self._h, None, _p64(off), _p32(idx), len(idx)))
        return ext.astype(bool), off, idx[:int(off[-1])]

    def autotune(self, reps=2):
        """Time every sweep configuration on the loaded image, configure the
        engine with the fastest; returns (best, [all points]) as dicts."""
        cap = 128
        allp = (Tuning * cap)()
        best = Tuning()
        n = ctypes.c_int32(0)
        _check(lib().pcx_batch_autotune(self._h, reps, ctypes.byref(best), allp, cap, ctypes.byref(n)))
        return best.as_dict(), [allp[i].as_dict() for i in range(min(n.value, cap))]


class Server:
    """Continuous-batching fixpoint server over registered PIR networks.

    submit() queues a store (int32 lb/ub arrays that stay owned by the
    caller and receive the fixpoint); step() runs one tick; poll() returns
    completed (tag, Result, latency_ms) tuples.  replay() drives a whole
    arrival trace in real time inside the library.
    """

    def __init__(self, **kw):
        c = ServerConfig()
        lib().pcx_server_config_default(ctypes.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        h = _VP()
        _check(lib().pcx_server_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.config = c
        self._nets = []
        self._live = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().pcx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_network(self, pir):
        n = ctypes.c_int32(-1)
        _check(lib().pcx_server_add_network(self._h, pir._h, ctypes.byref(n)))
        self._nets.append(pir)          # keeps the engine alive while registered
        return n.value

    def set_root(self, net, lb, ub):
        """The network's told store: later requests inside it start from
        their meet with its root fixpoint (pcx_server_set_root)."""
        nv = _store_pair(lb, ub)
        _check(lib().pcx_server_set_root(self._h, int(net), _p32(lb, nv), _p32(ub, nv), nv))

    def submit(self, net, lb, ub, tag):
        nv = _store_pair(lb, ub)
        _check(lib().pcx_server_submit(self._h, int(net), _p32(lb, nv), _p32(ub, nv), nv, int(tag)))
        self._live[int(tag)] = (lb, ub)   # the arrays must outlive the request

    def step(self):
        n = ctypes.c_int32(0)
        _check(lib().pcx_server_step(self._h, ctypes.byref(n)))
        return n.value

    def stats(self):
        st = ServerStats()
        _check(lib().pcx_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def poll(self, max_items=256):
        buf = (Served * max_items)()
        got = int(lib().pcx_server_poll(self._h, buf, max_items))
        out = []
        for i in range(got):
            self._live.pop(buf[i].tag, None)
            out.append((buf[i].tag, buf[i].result, buf[i].latency_ms))
        return out

    def replay(self, arrival_s, net, lb_all, ub_all, store_off, plain=False):
        """Serve a trace; returns (list of Served, wall seconds)."""
        n = len(arrival_s)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        net = np.ascontiguousarray(net, dtype=np.int32)
        store_off = np.ascontiguousarray(store_off, dtype=np.int64)
        if len(net) != n or len(store_off) != n + 1:
            raise ValueError("net needs one entry per arrival and store_off one more")
        if n and (lb_all.size < store_off[-1] or ub_all.size < store_off[-1]):
            raise ValueError("lb_all / ub_all are shorter than store_off[-1]")
        out = (Served * max(n, 1))()
        wall = ctypes.c_double(0.0)
        _check(lib().pcx_server_replay(self._h, n, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       _p32(net), _p32(lb_all), _p32(ub_all), _p64(store_off),
                                       1 if plain else 0, out, ctypes.byref(wall)))
        return list(out)[:n], wall.value


def default_tuning_path():
    """$PCX_TUNING, else ~/.cache/pcx/tuning.json."""
    p = os.environ.get("PCX_TUNING")
    if p:
        return p
    d = os.path.join(os.path.expanduser("~"), ".cache", "pcx")
    try:
        os.makedirs(d, exist_ok=True)
    except OSError:
        import tempfile
        d = tempfile.gettempdir()
    return os.path.join(d, "tuning.json")


def load_tuning(path, key):
    """A tuning saved by save_tuning for `key`, or None."""
    import json
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def save_tuning(path, key, best, points=None):
    """Persist an autotuning result as JSON under `key` (other keys kept)."""
    import json
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[key] = {"best": best, "points": points or []}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, path)


def apply_tuning(pir, best):
    """Configure the engine with a saved tuning; returns the mode to run."""
    pir.set_config(**{k: best[k] for k in ("props_per_wave", "lds_mode", "compact_tables",
                                           "stream_tables", "event_driven", "lockstep")
                      if k in best})
    return best["mode"]


# ---- synthetic workloads (host code, no GPU) ------------------------------
def gen_params(**kw):
    p = GenParams()
    lib().pcxg_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gen_network(p):
    """Returns (op, x, y, z, lb, ub, witness) in told order."""
    P, V = int(p.nprops), int(p.nvars)
    op, x, y, z = (np.zeros(P, np.int32) for _ in range(4))
    lb, ub, w = (np.zeros(V, np.int32) for _ in range(3))
    _check(lib().pcxg_network(ctypes.byref(p), _p32(op), _p32(x), _p32(y), _p32(z),
                              _p32(lb), _p32(ub), _p32(w)))
    return op, x, y, z, lb, ub, w


def gen_splits(seed, lb, ub, n_inst, kmin, kmax):
    V = _store_pair(lb, ub)
    lo = np.zeros(n_inst * V, np.int32)
    hi = np.zeros(n_inst * V, np.int32)
    _check(lib().pcxg_splits(seed, _p32(lb), _p32(ub), V, n_inst, kmin, kmax, _p32(lo), _p32(hi)))
    return lo, hi


def gen_bits_network(seed, nvars, nprops, kinds=0, dens_lo=0.3, dens_hi=1.0, planted=1,
                     witness=False):
    """Returns (op, x, y, z, words[, witness]) of a bitset network in told order."""
    op, x, y, z = (np.zeros(nprops, np.int32) for _ in range(4))
    w = np.zeros(nvars, np.uint64)
    wit = np.zeros(nvars, np.int32)
    _check(lib().pcxg_bits_network(seed, nvars, nprops, kinds, dens_lo, dens_hi, planted,
                                   _p32(op), _p32(x), _p32(y), _p32(z), _pu64(w), _p32(wit)))
    return (op, x, y, z, w, wit) if witness else (op, x, y, z, w)


def gen_bits_splits(seed, words, n_inst, kmin, kmax):
    out = np.zeros(n_inst * len(words), np.uint64)
    _check(lib().pcxg_bits_splits(seed, _pu64(words), len(words), n_inst, kmin, kmax, _pu64(out)))
    return out


def gen_subsets(seed, nprops, n_inst, mmin, mmax):
    off = np.zeros(n_inst + 1, np.int64)
    _check(lib().pcxg_subsets(seed, nprops, n_inst, mmin, mmax, _p64(off), None))
    idx = np.zeros(max(int(off[-1]), 1), np.int32)
    _check(lib().pcxg_subsets(seed, nprops, n_inst, mmin, mmax, _p64(off), _p32(idx)))
    return off, idx[:int(off[-1])]


class Model:
    """A FlatZinc-style model decomposed to PIR ternary form (include/pcx_model.h).

    Model(text) parses declarations and constraints; add(text) tells more, as
    the reference's tests tell further constraints into a PIR
    (tests/pir_test.cpp:577-581).  table(frm) gives the propagators created
    since `frm` (for PIR.tell), store() the declared and restricted domains,
    var(name) a variable's index.
    """

    def __init__(self, text):
        h = _VP()
        _check(lib().pcx_model_parse(text.encode(), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().pcx_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, text):
        _check(lib().pcx_model_add(self._h, text.encode()))

    def num_vars(self):
        return int(lib().pcx_model_num_vars(self._h))

    def num_props(self):
        return int(lib().pcx_model_num_props(self._h))

    def table(self, frm=0):
        n = self.num_props() - frm
        op, x, y, z = (np.zeros(max(n, 0), np.int32) for _ in range(4))
        _check(lib().pcx_model_table(self._h, frm, _p32(op), _p32(x), _p32(y), _p32(z)))
        return op, x, y, z

    def store(self):
        n = self.num_vars()
        lb, ub = np.zeros(n, np.int32), np.zeros(n, np.int32)
        _check(lib().pcx_model_store(self._h, _p32(lb), _p32(ub)))
        return lb, ub

    def var(self, name):
        return int(lib().pcx_model_var(self._h, name.encode()))

    def var_name(self, v):
        r = lib().pcx_model_var_name(self._h, v)
        return None if r is None else r.decode()

    def formulas(self):
        """PC's formula trees of the constraints told so far: (nodes, off),
        nodes an (n, 4) int32 array in postorder (kinds PCX_FT_* of
        pcx_model.h), formula f at nodes[off[f]:off[f + 1]]."""
        n = int(lib().pcx_model_num_formula_nodes(self._h))
        nf = int(lib().pcx_model_num_formulas(self._h))
        nodes = np.zeros((max(n, 1), 4), np.int32)
        off = np.zeros(nf + 1, np.int64)
        _check(lib().pcx_model_formulas(self._h, nodes.ctypes.data_as(_P32), _p64(off)))
        return nodes[:n], off

    def meet_store(self, lb, ub):
        """The current store met with the model's domains; variables created
        since the store was made are appended with their domains."""
        mlb, mub = self.store()
        n = len(lb)
        out_lb, out_ub = mlb.copy(), mub.copy()
        out_lb[:n] = np.maximum(lb, mlb[:n])
        out_ub[:n] = np.minimum(ub, mub[:n])
        return out_lb, out_ub


_FZN = {ADD: "int_plus({y}, {z}, {x})", MUL: "int_times({y}, {z}, {x})",
        TDIV: "int_tdiv({y}, {z}, {x})", FDIV: "int_fdiv({y}, {z}, {x})",
        CDIV: "int_cdiv({y}, {z}, {x})", EDIV: "int_ediv({y}, {z}, {x})",
        MIN: "int_min({y}, {z}, {x})", MAX: "int_max({y}, {z}, {x})",
        EQ: "int_eq_reif({y}, {z}, {x})", LEQ: "int_le_reif({y}, {z}, {x})"}


def flatzinc(op, x, y, z, lb, ub):
    """A ternary network as FlatZinc text (variables v0, v1, ...): each
    propagator x = y <op> z as the predicate that Model decomposes back into
    exactly that propagator."""
    out = [f"var {int(a)}..{int(b)}: v{i};" for i, (a, b) in enumerate(zip(lb, ub))]
    for o, a, b, c in zip(op.tolist(), x.tolist(), y.tolist(), z.tolist()):
        out.append("constraint " + _FZN[o].format(x=f"v{a}", y=f"v{b}", z=f"v{c}") + ";")
    return "\n".join(out)
