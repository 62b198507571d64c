"""Python front end of libBaryTree_gpu.so (ctypes over the C ABI).

Mirrors the reference's binding, interfaces/python/BaryTreeInterface.py:14-152
(the Kernel / Singularity / Approximation / ComputeType enums and
callTreedriver with the same arguments and defaults), so scripts written for
the reference run unchanged; GPUpresent is accepted and ignored because every
call runs on the GPU.  The extensions (precision choice, plans, direct sums,
statistics, batches of independent systems, the autotuner) are exposed as
solve(), Plan, analyze(), solve_batched(), Plan.batched() and autotune().

The library is loaded from this package directory.  Nothing here falls back
to a CPU path: a missing library raises at load time.
"""
import ctypes
import enum
import os

import numpy as np

LIB_NAME = "libBaryTree_gpu.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PRECISION_FP64 = 0
PRECISION_MIXED = 1


class Kernel(enum.IntEnum):
    NO_KERNEL = 0
    COULOMB = 1
    YUKAWA = 2
    REGULARIZED_COULOMB = 3
    REGULARIZED_YUKAWA = 4
    ATAN = 5
    TCF = 6
    DCF = 7
    SIN_OVER_R = 8
    MQ = 9
    RBS_U = 10
    RBS_V = 11
    USER = 12


class Singularity(enum.IntEnum):
    NO_SINGULARITY = 0
    SKIPPING = 1
    SUBTRACTION = 2


class Approximation(enum.IntEnum):
    NO_APPROX = 0
    LAGRANGE = 1
    HERMITE = 2


class ComputeType(enum.IntEnum):
    NO_COMPUTE_TYPE = 0
    PARTICLE_CLUSTER = 1
    CLUSTER_PARTICLE = 2
    CLUSTER_CLUSTER = 3


class Stats(ctypes.Structure):
    """BaryTreeStats of include/BaryTreeInterface.h."""
    _fields_ = [(n, ctypes.c_longlong) for n in (
        "num_pp", "num_pc", "num_cp", "num_cc",
        "pp_ptwise", "pc_ptwise", "cp_ptwise", "cc_ptwise",
        "pc_replaced", "cp_replaced", "cc_replaced")] + [
        (n, ctypes.c_int) for n in ("source_nodes", "target_nodes", "interp_degree",
                                    "pts_per_cluster")] + [
        ("theta", ctypes.c_double),
        ("max_per_source_leaf", ctypes.c_int), ("max_per_target_leaf", ctypes.c_int),
        ("near_tiles", ctypes.c_int), ("far_tiles", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in (
            "host_copy_s", "host_tree_s", "host_lists_s", "host_work_s", "h2d_s",
            "gpu_upward_s", "gpu_near_s", "gpu_far_s", "gpu_down_s", "gpu_finish_s",
            "gpu_total_s")] + [
        (n, ctypes.c_int) for n in ("num_systems", "targets_per_thread", "persistent")] + [
        ("num_far_pairs", ctypes.c_longlong), ("variable_degree", ctypes.c_int),
        ("mean_pair_degree", ctypes.c_double), ("far_ptwise_variable", ctypes.c_longlong),
        ("sym_ptwise", ctypes.c_longlong), ("gpu_near_sym_s", ctypes.c_double),
        ("sym_slots", ctypes.c_longlong), ("model_fallback_pairs", ctypes.c_longlong),
        ("sym_evals", ctypes.c_longlong)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


SIZE_CLASSES = 32


class Schedule(ctypes.Structure):
    """BaryTreeSchedule: leaf size per size class (2^c <= max(M, N) < 2^(c+1)),
    tile shape and launch form of a batched solve."""
    _fields_ = [("targets_per_thread", ctypes.c_int), ("persistent", ctypes.c_int),
                ("leaf", ctypes.c_int * SIZE_CLASSES), ("trials", ctypes.c_int),
                ("tune_seconds", ctypes.c_double)]

    @classmethod
    def make(cls, leaf_by_class, targets_per_thread=2, persistent=1):
        s = cls()
        s.targets_per_thread = targets_per_thread
        s.persistent = persistent
        for c, v in dict(leaf_by_class).items():
            s.leaf[int(c)] = int(v)
        return s

    def as_dict(self):
        return {"targets_per_thread": self.targets_per_thread, "persistent": self.persistent,
                "leaf": {c: self.leaf[c] for c in range(SIZE_CLASSES) if self.leaf[c] > 0},
                "trials": self.trials, "tune_seconds": self.tune_seconds}


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    """The loaded libBaryTree_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'`")
        l = ctypes.CDLL(LIB_PATH)
        l.BaryTree_GetLastError.restype = ctypes.c_char_p
        for f in ("BaryTree_Solve", "BaryTree_Analyze", "BaryTree_GetLastStats",
                  "BaryTreePlan_Create", "BaryTreePlan_Execute", "BaryTreePlan_GetStats",
                  "BaryTreePlan_DirectSample", "BaryTreePlan_DirectSampleNoAlloc", "BaryTreePlan_SetPrecision",
                  "BaryTreePlan_SetSymmetric", "BaryTree_SetBetaLeafSizes",
                  "BaryTree_SolveBatched", "BaryTreePlan_CreateBatched", "BaryTree_Autotune",
                  "BaryTree_AnalyzeBatched", "BaryTreePlan_NumPairs", "BaryTreePlan_PairInfo",
                  "BaryTreePlan_PairFeatures", "BaryTreePlan_LabelPairs",
                  "BaryTreePlan_SetPairDegrees", "BaryTreePlan_UseDegreeModel",
                  "BaryTree_LabelPairs", "BaryTree_LabelPairsSubtraction", "BaryTreeModel_Create", "BaryTreeModel_Load",
                  "BaryTreeModel_Save", "BaryTreeModel_GetParams", "BaryTreeModel_SetParams",
                  "BaryTreeModel_Gradient", "BaryTreeModel_Train", "BaryTreeModel_Predict",
                  "BaryTreeModel_GetFeatureRange", "BaryTreeModel_SetFeatureRange",
                  "BaryTree_SolveLearned", "BaryTree_AnalyzeDegrees", "BaryTree_SolveDevice",
                  "BaryTree_SolveLearnedDevice"):
            getattr(l, f).restype = ctypes.c_int
        l.BaryTreeInterface.restype = None
        l.BaryTreePlan_Destroy.restype = None
        l.BaryTreeModel_Destroy.restype = None
        l.BaryTreeModel_Create.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        l.BaryTreeModel_Train.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_ulonglong, ctypes.c_void_p]
        l.BaryTreeModel_Predict.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
        l.BaryTreePlan_UseDegreeModel.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]
        l.BaryTree_EmptyDeviceCache.restype = ctypes.c_ulonglong
        l.BaryTree_CachedDeviceBytes.restype = ctypes.c_ulonglong
        l.BaryTree_SetDeviceCacheLimit.restype = ctypes.c_int
        l.BaryTree_SetDeviceCacheLimit.argtypes = [ctypes.c_ulonglong]
        _lib = l
    return _lib


def last_error():
    return lib().BaryTree_GetLastError().decode()


def empty_device_cache():
    """Frees the device blocks that destroyed plans left in the library's
    cache; returns the bytes freed."""
    return int(lib().BaryTree_EmptyDeviceCache())


def set_device_cache_limit(nbytes):
    """Most bytes of released plan buffers the library keeps per GPU."""
    lib().BaryTree_SetDeviceCacheLimit(ctypes.c_ulonglong(int(nbytes)))


def cached_device_bytes():
    return int(lib().BaryTree_CachedDeviceBytes())


def _arr(a):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _kp(kernel_params):
    if kernel_params is None:
        return np.zeros(1)
    return np.ascontiguousarray(np.atleast_1d(np.asarray(kernel_params, dtype=np.float64)))


def solve(targets, sources, kernel, kernel_params, singularity, approximation, compute_type,
          theta, degree, max_source_leaf, max_target_leaf, size_check=1.0, beta=-1.0,
          precision=PRECISION_FP64, verbosity=0):
    """One BaryTree_Solve call.  targets = (x, y, z, q) (q may be None),
    sources = (x, y, z, q, w) (w may be None).  Returns the potential."""
    tx, ty, tz, tq = map(_arr, targets)
    sx, sy, sz, sq, sw = map(_arr, sources)
    kp = _kp(kernel_params)
    out = np.zeros(len(tx))
    rc = lib().BaryTree_Solve(
        ctypes.c_int(len(tx)), ctypes.c_int(len(sx)), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(tq),
        _ptr(sx), _ptr(sy), _ptr(sz), _ptr(sq), _ptr(sw), _ptr(out), ctypes.c_int(int(kernel)),
        ctypes.c_int(len(kp)), _ptr(kp), ctypes.c_int(int(singularity)),
        ctypes.c_int(int(approximation)), ctypes.c_int(int(compute_type)),
        ctypes.c_double(theta), ctypes.c_int(degree), ctypes.c_int(max_source_leaf),
        ctypes.c_int(max_target_leaf), ctypes.c_double(size_check), ctypes.c_double(beta),
        ctypes.c_int(verbosity), ctypes.c_int(precision))
    if rc != 0:
        raise RuntimeError(f"BaryTree_Solve failed: {last_error()}")
    return out


BETA_LEAVES_CPU, BETA_LEAVES_GPU = 0, 1


def set_beta_leaf_sizes(which):
    """Leaf sizes of the beta auto-tune for later calls on this thread:
    BETA_LEAVES_CPU (the reference CPU build's, default) or BETA_LEAVES_GPU
    (its GPU build's, 3000/3000)."""
    if lib().BaryTree_SetBetaLeafSizes(ctypes.c_int(int(which))) != 0:
        raise RuntimeError(f"BaryTree_SetBetaLeafSizes failed: {last_error()}")


def last_stats():
    s = Stats()
    if lib().BaryTree_GetLastStats(ctypes.byref(s)) != 0:
        return None
    return s.as_dict()


def analyze(targets, sources, kernel, kernel_params, singularity, approximation, compute_type,
            theta, degree, max_source_leaf, max_target_leaf, size_check=1.0, beta=-1.0):
    """Host-only tree and list construction; returns the interaction counts."""
    tx, ty, tz = map(_arr, targets[:3])
    sx, sy, sz = map(_arr, sources[:3])
    kp = _kp(kernel_params)
    s = Stats()
    rc = lib().BaryTree_Analyze(
        ctypes.c_int(len(tx)), ctypes.c_int(len(sx)), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(sx),
        _ptr(sy), _ptr(sz), ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
        ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
        ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
        ctypes.c_int(max_source_leaf), ctypes.c_int(max_target_leaf),
        ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.byref(s))
    if rc != 0:
        raise RuntimeError(f"BaryTree_Analyze failed: {last_error()}")
    return s.as_dict()


def analyze_degrees(targets, sources, kernel, kernel_params, singularity, approximation,
                    compute_type, theta, degree, max_source_leaf, max_target_leaf, degrees=None,
                    size_check=1.0, beta=-1.0):
    """BaryTree_AnalyzeDegrees (host only): (stats, features); with degrees,
    the variable-degree work is built and checked and the stats describe it."""
    tx, ty, tz = map(_arr, targets[:3])
    sx, sy, sz = map(_arr, sources[:3])
    kp = _kp(kernel_params)
    args = (ctypes.c_int(len(tx)), ctypes.c_int(len(sx)), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(sx),
            _ptr(sy), _ptr(sz), ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
            ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
            ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
            ctypes.c_int(max_source_leaf), ctypes.c_int(max_target_leaf),
            ctypes.c_double(size_check), ctypes.c_double(beta))
    s = Stats()
    if lib().BaryTree_AnalyzeDegrees(*args, 0, None, None, ctypes.byref(s)) != 0:
        raise RuntimeError(f"BaryTree_AnalyzeDegrees failed: {last_error()}")
    n = s.num_far_pairs
    f = np.zeros((n, 8), dtype=np.float32)
    d = None if degrees is None else np.ascontiguousarray(degrees, dtype=np.int32)
    rc = lib().BaryTree_AnalyzeDegrees(*args, ctypes.c_int(n), None if d is None else
                                       d.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                       f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       ctypes.byref(s))
    if rc != 0:
        raise RuntimeError(f"BaryTree_AnalyzeDegrees failed: {last_error()}")
    return s.as_dict(), f


def _batch(systems):
    """Concatenates [(targets, sources), ...] into offset arrays and particle
    arrays (targets = (x, y, z, q), sources = (x, y, z, q, w))."""
    toffs, soffs = [0], [0]
    cols = [[] for _ in range(9)]
    for tgt, src in systems:
        tx, ty, tz, tq = tgt
        sx, sy, sz, sq, sw = src
        m, n = len(tx), len(sx)
        toffs.append(toffs[-1] + m)
        soffs.append(soffs[-1] + n)
        for k, a in enumerate((tx, ty, tz, tq if tq is not None else np.zeros(m),
                               sx, sy, sz, sq, sw if sw is not None else np.ones(n))):
            cols[k].append(np.asarray(a, dtype=np.float64))
    arrays = [np.ascontiguousarray(np.concatenate(c)) if c else np.zeros(0) for c in cols]
    return (np.asarray(toffs, dtype=np.int32), np.asarray(soffs, dtype=np.int32), arrays)


class Batch:
    """Systems [(targets, sources), ...] concatenated once into the arrays
    BaryTree_SolveBatched takes (offsets, then targets x, y, z, q and sources
    x, y, z, q, w): a caller that solves the same batch layout again passes
    a Batch wherever a list of systems is accepted and skips the copies."""

    def __init__(self, systems):
        self.toffs, self.soffs, self.arrays = _batch(systems)

    def __len__(self):
        return len(self.toffs) - 1


def _batched(systems):
    if isinstance(systems, Batch):
        return systems.toffs, systems.soffs, systems.arrays
    return _batch(systems)


def _ints(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)))
    return a


def _iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)


def solve_batched(systems, kernel, kernel_params, singularity, approximation, compute_type,
                  theta, degree, max_source_leaf=None, max_target_leaf=None, schedule=None,
                  size_check=1.0, beta=-1.0, precision=PRECISION_FP64, verbosity=0):
    """BaryTree_SolveBatched over systems = [(targets, sources), ...] (or a
    Batch).  Leaf
    sizes: per-system arrays (or one int for all), else the schedule, else the
    autotuner.  Returns one potential array per system."""
    toffs, soffs, arr = _batched(systems)
    ns = len(toffs) - 1
    kp = _kp(kernel_params)
    msl, mtl = _ints(max_source_leaf, ns), _ints(max_target_leaf, ns)
    out = np.zeros(int(toffs[-1]))
    rc = lib().BaryTree_SolveBatched(
        ctypes.c_int(ns), _iptr(toffs), _iptr(soffs), *[_ptr(a) for a in arr], _ptr(out),
        ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
        ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
        ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
        _iptr(msl), _iptr(mtl), None if schedule is None else ctypes.byref(schedule),
        ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.c_int(verbosity),
        ctypes.c_int(precision))
    if rc != 0:
        raise RuntimeError(f"BaryTree_SolveBatched failed: {last_error()}")
    return [out[toffs[i]:toffs[i + 1]] for i in range(ns)]


def analyze_batched(systems, kernel, kernel_params, singularity, approximation, compute_type,
                    theta, degree, max_source_leaf=None, max_target_leaf=None, schedule=None,
                    size_check=1.0, beta=-1.0):
    """BaryTree_AnalyzeBatched: host-only trees and lists, counts summed over
    the systems (no GPU work)."""
    toffs, soffs, arr = _batched(systems)
    ns = len(toffs) - 1
    kp = _kp(kernel_params)
    msl, mtl = _ints(max_source_leaf, ns), _ints(max_target_leaf, ns)
    s = Stats()
    tx, ty, tz, _, sx, sy, sz, _, _ = arr
    rc = lib().BaryTree_AnalyzeBatched(
        ctypes.c_int(ns), _iptr(toffs), _iptr(soffs), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(sx),
        _ptr(sy), _ptr(sz), ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
        ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
        ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
        _iptr(msl), _iptr(mtl), None if schedule is None else ctypes.byref(schedule),
        ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.byref(s))
    if rc != 0:
        raise RuntimeError(f"BaryTree_AnalyzeBatched failed: {last_error()}")
    return s.as_dict()


def autotune(systems, kernel, kernel_params, singularity, approximation, compute_type, theta,
             degree, size_check=1.0, beta=-1.0, precision=PRECISION_FP64):
    """BaryTree_Autotune: returns the Schedule picked for this batch."""
    toffs, soffs, arr = _batched(systems)
    kp = _kp(kernel_params)
    sc = Schedule()
    rc = lib().BaryTree_Autotune(
        ctypes.c_int(len(systems)), _iptr(toffs), _iptr(soffs), *[_ptr(a) for a in arr],
        ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
        ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
        ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
        ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.c_int(precision),
        ctypes.byref(sc))
    if rc != 0:
        raise RuntimeError(f"BaryTree_Autotune failed: {last_error()}")
    return sc


class Plan:
    """Host setup once, GPU execution many times (BaryTreePlan_*)."""

    @classmethod
    def batched(cls, systems, kernel, kernel_params, singularity, approximation, compute_type,
                theta, degree, max_source_leaf=None, max_target_leaf=None, schedule=None,
                size_check=1.0, beta=-1.0, precision=PRECISION_FP64):
        """BaryTreePlan_CreateBatched; execute() returns the concatenated
        potential (split with .offsets)."""
        toffs, soffs, arr = _batched(systems)
        ns = len(toffs) - 1
        kp = _kp(kernel_params)
        msl, mtl = _ints(max_source_leaf, ns), _ints(max_target_leaf, ns)
        self = cls.__new__(cls)
        self.n_targets = int(toffs[-1])
        self.offsets = toffs
        self._h = ctypes.c_void_p()
        rc = lib().BaryTreePlan_CreateBatched(
            ctypes.byref(self._h), ctypes.c_int(ns), _iptr(toffs), _iptr(soffs),
            *[_ptr(a) for a in arr], ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
            ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
            ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
            _iptr(msl), _iptr(mtl), None if schedule is None else ctypes.byref(schedule),
            ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.c_int(precision))
        if rc != 0:
            self._h = None
            raise RuntimeError(f"BaryTreePlan_CreateBatched failed: {last_error()}")
        return self

    def __init__(self, targets, sources, kernel, kernel_params, singularity, approximation,
                 compute_type, theta, degree, max_source_leaf, max_target_leaf, size_check=1.0,
                 beta=-1.0, precision=PRECISION_FP64):
        tx, ty, tz, tq = map(_arr, targets)
        sx, sy, sz, sq, sw = map(_arr, sources)
        kp = _kp(kernel_params)
        self.n_targets = len(tx)
        self.offsets = np.asarray([0, len(tx)], dtype=np.int32)
        self._h = ctypes.c_void_p()
        rc = lib().BaryTreePlan_Create(
            ctypes.byref(self._h), ctypes.c_int(len(tx)), ctypes.c_int(len(sx)), _ptr(tx),
            _ptr(ty), _ptr(tz), _ptr(tq), _ptr(sx), _ptr(sy), _ptr(sz), _ptr(sq), _ptr(sw),
            ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp),
            ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
            ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
            ctypes.c_int(max_source_leaf), ctypes.c_int(max_target_leaf),
            ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.c_int(precision))
        if rc != 0:
            self._h = None
            raise RuntimeError(f"BaryTreePlan_Create failed: {last_error()}")

    def execute(self, download=True):
        out = np.zeros(self.n_targets) if download else None
        rc = lib().BaryTreePlan_Execute(self._h, _ptr(out))
        if rc != 0:
            raise RuntimeError(f"BaryTreePlan_Execute failed: {last_error()}")
        return out

    def set_precision(self, precision):
        if lib().BaryTreePlan_SetPrecision(self._h, ctypes.c_int(precision)) != 0:
            raise RuntimeError(f"BaryTreePlan_SetPrecision failed: {last_error()}")

    def set_symmetric(self, enable):
        """Symmetric near field of later executions on (default) or off (one
        evaluation per ordered pair, as the reference)."""
        if lib().BaryTreePlan_SetSymmetric(self._h, ctypes.c_int(1 if enable else 0)) != 0:
            raise RuntimeError(f"BaryTreePlan_SetSymmetric failed: {last_error()}")

    def stats(self):
        s = Stats()
        lib().BaryTreePlan_GetStats(self._h, ctypes.byref(s))
        return s.as_dict()

    def direct_sample(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros(len(idx))
        rc = lib().BaryTreePlan_DirectSample(self._h, ctypes.c_int(len(idx)),
                                             idx.ctypes.data_as(_ip), _ptr(out))
        if rc != 0:
            raise RuntimeError(f"BaryTreePlan_DirectSample failed: {last_error()}")
        return out

    # ---- far-field pairs and their degrees
    def num_pairs(self):
        return int(lib().BaryTreePlan_NumPairs(self._h))

    def pair_info(self):
        """(kind, system) arrays: kind 1 PC, 2 CP, 3 CC."""
        n = self.num_pairs()
        kind = np.zeros(n, dtype=np.int32)
        system = np.zeros(n, dtype=np.int32)
        lib().BaryTreePlan_PairInfo(self._h, _iptr(kind), _iptr(system))
        return kind, system

    def pair_features(self):
        n = self.num_pairs()
        f = np.zeros((n, PAIR_FEATURES), dtype=np.float32)
        if n and lib().BaryTreePlan_PairFeatures(self._h, f.ctypes.data_as(_fp)) != 0:
            raise RuntimeError(f"BaryTreePlan_PairFeatures failed: {last_error()}")
        return f

    def label_pairs(self, tol, first=0, count=None, with_errors=False):
        n = self.num_pairs()
        count = n - first if count is None else count
        labels = np.zeros(count, dtype=np.int32)
        errors = np.zeros((count, LABEL_DEGREES)) if with_errors else None
        rc = lib().BaryTreePlan_LabelPairs(self._h, ctypes.c_int(first), ctypes.c_int(count),
                                           ctypes.c_double(tol), _iptr(labels), _ptr(errors))
        if rc != 0:
            raise RuntimeError(f"BaryTreePlan_LabelPairs failed: {last_error()}")
        return (labels, errors) if with_errors else labels

    def set_pair_degrees(self, degrees):
        """One degree per far-field pair; None restores the fixed degree."""
        d = None if degrees is None else np.ascontiguousarray(degrees, dtype=np.int32)
        if d is not None and len(d) != self.num_pairs():
            raise ValueError("need one degree per far-field pair")
        if lib().BaryTreePlan_SetPairDegrees(self._h, _iptr(d)) != 0:
            raise RuntimeError(f"BaryTreePlan_SetPairDegrees failed: {last_error()}")

    def use_model(self, model, delta=0.5):
        if lib().BaryTreePlan_UseDegreeModel(self._h, model._h, ctypes.c_float(delta)) != 0:
            raise RuntimeError(f"BaryTreePlan_UseDegreeModel failed: {last_error()}")

    def close(self):
        if self._h:
            lib().BaryTreePlan_Destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PAIR_FEATURES = 8
LABEL_DEGREES = 11
MODEL_PARAMS = 9546
_fp = ctypes.POINTER(ctypes.c_float)


def label_pairs(kind, tbeg, tcnt, sbeg, scnt, tbox, sbox, targets, sources, kernel, kernel_param,
                tol):
    """BaryTree_LabelPairs on explicitly given pairs: returns (labels, errors)."""
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (kind, tbeg, tcnt, sbeg, scnt)]
    tbox, sbox = (np.ascontiguousarray(a, dtype=np.float64) for a in (tbox, sbox))
    tx, ty, tz = map(_arr, targets[:3])
    sx, sy, sz, sq = map(_arr, sources[:4])
    n = len(ints[0])
    labels = np.zeros(n, dtype=np.int32)
    errors = np.zeros((n, LABEL_DEGREES))
    rc = lib().BaryTree_LabelPairs(
        ctypes.c_int(n), *[_iptr(a) for a in ints], _ptr(tbox), _ptr(sbox), ctypes.c_int(len(tx)),
        _ptr(tx), _ptr(ty), _ptr(tz), ctypes.c_int(len(sx)), _ptr(sx), _ptr(sy), _ptr(sz),
        _ptr(sq), ctypes.c_int(int(kernel)), ctypes.c_double(kernel_param), ctypes.c_double(tol),
        _iptr(labels), _ptr(errors))
    if rc != 0:
        raise RuntimeError(f"BaryTree_LabelPairs failed: {last_error()}")
    return labels, errors


def label_pairs_subtraction(kind, tbeg, tcnt, sbeg, scnt, tbox, sbox, targets, sources, kernel,
                            kernel_params, tol):
    """BaryTree_LabelPairsSubtraction: targets = (x, y, z, q), sources =
    (x, y, z, q, w); kernel_params {kappa} (Coulomb, Yukawa) or {alpha,
    epsilon} (regularized Coulomb).  Returns (labels, errors)."""
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (kind, tbeg, tcnt, sbeg, scnt)]
    tbox, sbox = (np.ascontiguousarray(a, dtype=np.float64) for a in (tbox, sbox))
    tx, ty, tz, tq = map(_arr, targets[:4])
    sx, sy, sz, sq, sw = map(_arr, sources[:5])
    kp = _kp(kernel_params)
    n = len(ints[0])
    labels = np.zeros(n, dtype=np.int32)
    errors = np.zeros((n, LABEL_DEGREES))
    rc = lib().BaryTree_LabelPairsSubtraction(
        ctypes.c_int(n), *[_iptr(a) for a in ints], _ptr(tbox), _ptr(sbox), ctypes.c_int(len(tx)),
        _ptr(tx), _ptr(ty), _ptr(tz), _ptr(tq), ctypes.c_int(len(sx)), _ptr(sx), _ptr(sy), _ptr(sz),
        _ptr(sq), _ptr(sw), ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp), ctypes.c_double(tol),
        _iptr(labels), _ptr(errors))
    if rc != 0:
        raise RuntimeError(f"BaryTree_LabelPairsSubtraction failed: {last_error()}")
    return labels, errors


class DegreeModel:
    """BaryTreeDegreeModel: the MLP that picks far-field degrees."""

    def __init__(self, seed=0, path=None):
        self._h = ctypes.c_void_p()
        if path is not None:
            rc = lib().BaryTreeModel_Load(ctypes.byref(self._h), str(path).encode())
        else:
            rc = lib().BaryTreeModel_Create(ctypes.byref(self._h), ctypes.c_ulonglong(seed))
        if rc != 0:
            self._h = None
            raise RuntimeError(f"degree model: {last_error()}")

    def save(self, path):
        if lib().BaryTreeModel_Save(self._h, str(path).encode()) != 0:
            raise RuntimeError(f"BaryTreeModel_Save failed: {last_error()}")

    def get_params(self):
        p = np.zeros(MODEL_PARAMS, dtype=np.float32)
        m = np.zeros(PAIR_FEATURES, dtype=np.float32)
        sc = np.zeros(PAIR_FEATURES, dtype=np.float32)
        lib().BaryTreeModel_GetParams(self._h, p.ctypes.data_as(_fp), m.ctypes.data_as(_fp),
                                      sc.ctypes.data_as(_fp))
        return p, m, sc

    def set_params(self, params=None, mean=None, scale=None):
        conv = [None if a is None else np.ascontiguousarray(a, dtype=np.float32)
                for a in (params, mean, scale)]
        lib().BaryTreeModel_SetParams(self._h, *[None if a is None else a.ctypes.data_as(_fp)
                                                 for a in conv])

    def feature_range(self):
        """(lo, hi) of every feature over the training data, or None."""
        lo = np.zeros(PAIR_FEATURES, dtype=np.float32)
        hi = np.zeros(PAIR_FEATURES, dtype=np.float32)
        rc = lib().BaryTreeModel_GetFeatureRange(self._h, lo.ctypes.data_as(_fp), hi.ctypes.data_as(_fp))
        if rc < 0:
            raise RuntimeError(f"BaryTreeModel_GetFeatureRange failed: {last_error()}")
        return (lo, hi) if rc == 1 else None

    def set_feature_range(self, lo, hi):
        """Sets the training range (None removes it and the guard)."""
        if lo is None or hi is None:
            rc = lib().BaryTreeModel_SetFeatureRange(self._h, None, None)
        else:
            lo = np.ascontiguousarray(lo, dtype=np.float32)
            hi = np.ascontiguousarray(hi, dtype=np.float32)
            rc = lib().BaryTreeModel_SetFeatureRange(self._h, lo.ctypes.data_as(_fp), hi.ctypes.data_as(_fp))
        if rc != 0:
            raise RuntimeError(f"BaryTreeModel_SetFeatureRange failed: {last_error()}")

    def gradient(self, features, labels):
        x = np.ascontiguousarray(features, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        g = np.zeros(MODEL_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        rc = lib().BaryTreeModel_Gradient(self._h, ctypes.c_int(len(y)), x.ctypes.data_as(_fp),
                                          _iptr(y), g.ctypes.data_as(_fp), ctypes.byref(loss))
        if rc != 0:
            raise RuntimeError(f"BaryTreeModel_Gradient failed: {last_error()}")
        return loss.value, g

    def train(self, features, labels, epochs=20, batch=4096, lr=1e-3, seed=0):
        x = np.ascontiguousarray(features, dtype=np.float32)
        y = np.ascontiguousarray(labels, dtype=np.int32)
        el = np.zeros(epochs, dtype=np.float32)
        rc = lib().BaryTreeModel_Train(self._h, len(y), x.ctypes.data, y.ctypes.data, epochs,
                                       batch, lr, seed, el.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"BaryTreeModel_Train failed: {last_error()}")
        return el

    def predict(self, features, delta=0.5, with_logits=False):
        x = np.ascontiguousarray(features, dtype=np.float32)
        n = x.shape[0]
        d = np.zeros(n, dtype=np.int32)
        z = np.zeros((n, 10), dtype=np.float32) if with_logits else None
        rc = lib().BaryTreeModel_Predict(self._h, n, x.ctypes.data, delta, d.ctypes.data,
                                         None if z is None else z.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"BaryTreeModel_Predict failed: {last_error()}")
        return (d, z) if with_logits else d

    def close(self):
        if self._h:
            lib().BaryTreeModel_Destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_learned(targets, sources, kernel, kernel_params, singularity, approximation,
                  compute_type, theta, degree, max_source_leaf, max_target_leaf, model=None,
                  delta=0.5, size_check=1.0, beta=-1.0, precision=PRECISION_FP64, verbosity=0):
    """BaryTree_SolveLearned: far-field degrees from `model` (None: fixed)."""
    tx, ty, tz, tq = map(_arr, targets)
    sx, sy, sz, sq, sw = map(_arr, sources)
    kp = _kp(kernel_params)
    out = np.zeros(len(tx))
    rc = lib().BaryTree_SolveLearned(
        ctypes.c_int(len(tx)), ctypes.c_int(len(sx)), _ptr(tx), _ptr(ty), _ptr(tz), _ptr(tq),
        _ptr(sx), _ptr(sy), _ptr(sz), _ptr(sq), _ptr(sw), _ptr(out), ctypes.c_int(int(kernel)),
        ctypes.c_int(len(kp)), _ptr(kp), ctypes.c_int(int(singularity)),
        ctypes.c_int(int(approximation)), ctypes.c_int(int(compute_type)),
        ctypes.c_double(theta), ctypes.c_int(degree), ctypes.c_int(max_source_leaf),
        ctypes.c_int(max_target_leaf), ctypes.c_double(size_check), ctypes.c_double(beta),
        ctypes.c_int(verbosity), ctypes.c_int(precision),
        None if model is None else model._h, ctypes.c_float(delta))
    if rc != 0:
        raise RuntimeError(f"BaryTree_SolveLearned failed: {last_error()}")
    return out


class DeviceBuffer:
    """float64 device memory allocated with the HIP runtime the library
    itself runs on (for callers without torch): from_numpy, to_numpy, free.
    Any object with data_ptr() and numel() (a torch CUDA tensor) serves the
    *_device calls as well."""

    def __init__(self, n):
        self._hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        if self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(max(self.n, 1) * 8)) != 0:
            raise RuntimeError("hipMalloc failed")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = cls(len(a))
        if b._hip.hipMemcpy(b.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy failed")
        return b

    def to_numpy(self):
        a = np.empty(self.n)
        if self._hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(a.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy failed")
        return a

    def data_ptr(self):
        return self.ptr.value

    def numel(self):
        return self.n

    def free(self):
        if self.ptr:
            self._hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _dev_ptr(a):
    """The device address of a device array (None passes through)."""
    return None if a is None else ctypes.c_void_p(a.data_ptr())


def solve_device(targets, sources, out, kernel, kernel_params, singularity, approximation, compute_type,
                 theta, degree, max_source_leaf, max_target_leaf, size_check=1.0, beta=-1.0,
                 precision=PRECISION_FP64, verbosity=0, model=None, delta=0.5):
    """BaryTree_SolveDevice (model None) or BaryTree_SolveLearnedDevice:
    targets = (x, y, z, q), sources = (x, y, z, q, w) and out are device
    arrays -- DeviceBuffers, or contiguous float64 torch CUDA tensors when
    torch's HIP runtime is the process's (torch initialised first) -- q, w
    may be None; the potential is written to out."""
    arrs = list(targets[:4]) + list(sources[:5])
    for a in arrs + [out]:
        if a is None or isinstance(a, DeviceBuffer):
            continue
        if (not getattr(a, "is_cuda", False) or not a.is_contiguous() or str(a.dtype) != "torch.float64"):
            raise ValueError("device arrays must be DeviceBuffers or contiguous float64 CUDA tensors")
    kp = _kp(kernel_params)
    args = [ctypes.c_int(int(targets[0].numel())), ctypes.c_int(int(sources[0].numel()))]
    args += [_dev_ptr(a) for a in arrs] + [_dev_ptr(out)]
    args += [ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _ptr(kp), ctypes.c_int(int(singularity)),
             ctypes.c_int(int(approximation)), ctypes.c_int(int(compute_type)), ctypes.c_double(theta),
             ctypes.c_int(degree), ctypes.c_int(max_source_leaf), ctypes.c_int(max_target_leaf),
             ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.c_int(verbosity), ctypes.c_int(precision)]
    if model is None:
        rc = lib().BaryTree_SolveDevice(*args)
        name = "BaryTree_SolveDevice"
    else:
        rc = lib().BaryTree_SolveLearnedDevice(*args, model._h, ctypes.c_float(delta))
        name = "BaryTree_SolveLearnedDevice"
    if rc != 0:
        raise RuntimeError(f"{name} failed: {last_error()}")
    return out


def callTreedriver(numTargets, numSources,
                   targetX, targetY, targetZ, targetValue,
                   sourceX, sourceY, sourceZ, sourceValue, sourceWeight,
                   kernelName, numberOfKernelParameters, kernelParameters, singularityHandling,
                   approximationName, computeType, GPUpresent, verbosity,
                   beta=None, theta=None, degree=None, sourceLeafSize=None, targetLeafSize=None,
                   sizeCheck=None):
    """Same arguments, defaults and result as the reference's callTreedriver
    (interfaces/python/BaryTreeInterface.py:82-152): either beta, or theta,
    degree and both leaf sizes; sizeCheck defaults to 1 (Lagrange) or 4
    (Hermite).  Returns the potential at the targets."""
    if beta:
        theta, degree, sourceLeafSize, targetLeafSize = -1.0, -1, 1, 1
    else:
        assert theta is not None, "If beta isn't provided, theta must be."
        assert degree is not None, "If beta isn't provided, degree must be."
        assert sourceLeafSize is not None, "If beta isn't provided, sourceLeafSize must be."
        assert targetLeafSize is not None, "If beta isn't provided, targetLeafSize must be."
        beta = -1
    if not sizeCheck:
        sizeCheck = 4.0 if approximationName == Approximation.HERMITE else 1.0
    kp = _kp(kernelParameters)
    arrays = [_arr(a) for a in (targetX, targetY, targetZ, targetValue, sourceX, sourceY,
                                sourceZ, sourceValue, sourceWeight)]
    result = np.zeros(numTargets)
    lib().BaryTreeInterface(
        ctypes.c_int(numTargets), ctypes.c_int(numSources), *[_ptr(a) for a in arrays],
        _ptr(result), ctypes.c_int(int(kernelName)), ctypes.c_int(numberOfKernelParameters),
        _ptr(kp), ctypes.c_int(int(singularityHandling)), ctypes.c_int(int(approximationName)),
        ctypes.c_int(int(computeType)), ctypes.c_double(theta), ctypes.c_int(degree),
        ctypes.c_int(sourceLeafSize), ctypes.c_int(targetLeafSize), ctypes.c_double(sizeCheck),
        ctypes.c_double(beta), ctypes.c_int(verbosity))
    return result
