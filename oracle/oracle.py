"""ctypes front end of the CPU oracle (libbt_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product package (barytree_amd) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbt_oracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libbt_refparts.so")

_dp = ctypes.POINTER(ctypes.c_double)


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in (
        "num_pp", "num_pc", "num_cp", "num_cc",
        "pp_ptwise", "pc_ptwise", "cp_ptwise", "cc_ptwise",
        "pc_replaced", "cp_replaced", "cc_replaced")] + [
        (n, ctypes.c_int) for n in (
            "source_nodes", "source_leaves", "target_nodes", "target_leaves", "batches",
            "max_depth_source", "max_depth_target", "interp_degree", "pts_per_cluster")] + [
        ("theta", ctypes.c_double),
        ("max_per_source_leaf", ctypes.c_int), ("max_per_target_leaf", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in (
            "t_setup", "t_upward", "t_lists", "t_compute", "t_downpass", "t_finish")] + [
        ("near_done", ctypes.c_longlong), ("far_done", ctypes.c_longlong)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


def build(ref=False):
    """Compile the oracle (and, with ref=True, the reference-parts library)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"] + (["ref"] if ref else []), check=True)


_LIBS = {}


def _load(path):
    if path not in _LIBS:
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        lib = ctypes.CDLL(path)
        _LIBS[path] = lib
    return _LIBS[path]


def lib():
    l = _load(LIB_PATH)
    l.orc_treecode.restype = ctypes.c_int
    l.orc_direct.restype = ctypes.c_int
    return l


def ref_available():
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    l = _load(REF_LIB_PATH)
    l.ref_treecode.restype = ctypes.c_int
    return l


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _call_tree(fn, tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, kernel_params, singularity,
               approximation, compute_type, theta, degree, max_source_leaf, max_target_leaf,
               size_check, beta):
    tx, ty, tz, tq, sx, sy, sz, sq, sw = map(_c, (tx, ty, tz, tq, sx, sy, sz, sq, sw))
    kp = np.ascontiguousarray(np.atleast_1d(np.asarray(kernel_params, dtype=np.float64)))
    out = np.zeros(len(tx))
    rc = fn(ctypes.c_int(len(tx)), ctypes.c_int(len(sx)),
            _p(tx), _p(ty), _p(tz), _p(tq), _p(sx), _p(sy), _p(sz), _p(sq), _p(sw), _p(out),
            ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)), _p(kp),
            ctypes.c_int(int(singularity)), ctypes.c_int(int(approximation)),
            ctypes.c_int(int(compute_type)), ctypes.c_double(theta), ctypes.c_int(degree),
            ctypes.c_int(max_source_leaf), ctypes.c_int(max_target_leaf),
            ctypes.c_double(size_check), ctypes.c_double(beta), ctypes.c_int(0))
    if rc != 0:
        raise RuntimeError(f"oracle returned {rc}")
    return out


def treecode(tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, kernel_params, singularity,
             approximation, compute_type, theta, degree, max_source_leaf, max_target_leaf,
             size_check=1.0, beta=-1.0):
    """BaryTreeInterface semantics on the CPU (FP64, serial)."""
    return _call_tree(lib().orc_treecode, tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel,
                      kernel_params, singularity, approximation, compute_type, theta, degree,
                      max_source_leaf, max_target_leaf, size_check, beta)


def ref_treecode(tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, kernel_params, singularity,
                 approximation, compute_type, theta, degree, max_source_leaf, max_target_leaf,
                 size_check=1.0, beta=-1.0):
    """Same call through the reference's own compiled sources (oracle/_ref)."""
    return _call_tree(ref_lib().ref_treecode, tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel,
                      kernel_params, singularity, approximation, compute_type, theta, degree,
                      max_source_leaf, max_target_leaf, size_check, beta)


def direct(tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, kernel_params, singularity):
    """directdriver semantics: FP64 direct sum plus subtraction correction."""
    tx, ty, tz, tq, sx, sy, sz, sq, sw = map(_c, (tx, ty, tz, tq, sx, sy, sz, sq, sw))
    kp = np.ascontiguousarray(np.atleast_1d(np.asarray(kernel_params, dtype=np.float64)))
    out = np.zeros(len(tx))
    rc = lib().orc_direct(ctypes.c_int(len(tx)), ctypes.c_int(len(sx)),
                          _p(tx), _p(ty), _p(tz), _p(tq), _p(sx), _p(sy), _p(sz), _p(sq),
                          _p(sw), _p(out), ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)),
                          _p(kp), ctypes.c_int(int(singularity)))
    if rc != 0:
        raise RuntimeError(f"oracle direct returned {rc}")
    return out


def last_stats():
    s = Stats()
    lib().orc_last_stats(ctypes.byref(s))
    return s.as_dict()


def set_sample_stride(stride):
    lib().orc_set_sample_stride(ctypes.c_int(int(stride)))


def set_threads(n):
    """Threads of the oracle's parallel loops (also those the reference-parts
    library runs through the restatement); results do not depend on it."""
    lib().orc_set_threads(ctypes.c_int(int(n)))
    if ref_available():
        ref_lib().orc_set_threads(ctypes.c_int(int(n)))


def max_threads():
    return int(lib().orc_max_threads())


class RefTiming(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "t_setup", "t_upward", "t_lists", "t_compute", "t_downpass", "t_finish")] + [
        (n, ctypes.c_longlong) for n in ("near_done", "far_done", "near_total", "far_total")]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


def ref_set_sample_stride(stride):
    """oracle/_ref: compute only every stride-th key's interaction lists."""
    ref_lib().ref_set_sample_stride(ctypes.c_int(int(stride)))


def ref_last_timing():
    """Phase times and point-wise counts (computed and full) of the last
    ref_treecode call."""
    t = RefTiming()
    ref_lib().ref_last_timing(ctypes.byref(t))
    return t.as_dict()


def source_clusters(x, y, z, q, w, degree, max_leaf, singularity, approximation=1):
    """Upward pass alone: (cx, cy, cz, Q, W) per node; coordinates (numnodes,
    P), charges (numnodes, P) for Lagrange, (numnodes, 8, P) for Hermite."""
    x, y, z, q, w = map(_c, (x, y, z, q, w))
    l = lib()
    nn = l.orc_source_clusters_apx(len(x), _p(x), _p(y), _p(z), _p(q), _p(w), degree, max_leaf,
                                   singularity, approximation, None, None, None, None, None)
    if nn < 0:
        raise RuntimeError(f"oracle returned {nn}")
    P = (degree + 1) ** 3
    planes = 8 if approximation == 2 else 1
    bufs = [np.zeros(nn * P) for _ in range(3)] + [np.zeros(nn * P * planes) for _ in range(2)]
    nn2 = l.orc_source_clusters_apx(len(x), _p(x), _p(y), _p(z), _p(q), _p(w), degree, max_leaf,
                                    singularity, approximation, *[_p(b) for b in bufs])
    assert nn2 == nn
    out = [b.reshape(nn, P) for b in bufs[:3]]
    shape = (nn, 8, P) if planes == 8 else (nn, P)
    return out + [b.reshape(shape) for b in bufs[3:]]


_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)


def label_pairs(kind, tbeg, tcnt, sbeg, scnt, tbox, sbox, tx, ty, tz, sx, sy, sz, sq, kernel, p0,
                tol):
    """Pair labels and per-degree errors (degrees 2..12), CPU FP64."""
    l = lib()
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (kind, tbeg, tcnt, sbeg, scnt)]
    dbl = [np.ascontiguousarray(a, dtype=np.float64) for a in (tbox, sbox, tx, ty, tz, sx, sy, sz, sq)]
    n = len(ints[0])
    labels = np.zeros(n, dtype=np.int32)
    errors = np.zeros((n, 11))
    l.orc_label_pairs.restype = ctypes.c_int
    rc = l.orc_label_pairs(ctypes.c_int(n), *[a.ctypes.data_as(_ip) for a in ints],
                           *[a.ctypes.data_as(_dp) for a in dbl], ctypes.c_int(int(kernel)),
                           ctypes.c_double(p0), ctypes.c_double(tol), labels.ctypes.data_as(_ip),
                           errors.ctypes.data_as(_dp))
    if rc != 0:
        raise RuntimeError("oracle label_pairs failed")
    return labels, errors


def label_pairs_ss(kind, tbeg, tcnt, sbeg, scnt, tbox, sbox, tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, p0,
                   p1, tol):
    """Pair labels of the subtraction kernel (both terms): target charges tq,
    source weights sw, p1 = 1/kappa^2 (Coulomb) or 1/alpha^2 (regularized
    Coulomb, p0 = epsilon); Yukawa ignores p1."""
    l = lib()
    ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (kind, tbeg, tcnt, sbeg, scnt)]
    dbl = [np.ascontiguousarray(a, dtype=np.float64) for a in (tbox, sbox, tx, ty, tz, tq, sx, sy, sz, sq, sw)]
    n = len(ints[0])
    labels = np.zeros(n, dtype=np.int32)
    errors = np.zeros((n, 11))
    l.orc_label_pairs_ss.restype = ctypes.c_int
    rc = l.orc_label_pairs_ss(ctypes.c_int(n), *[a.ctypes.data_as(_ip) for a in ints],
                              *[a.ctypes.data_as(_dp) for a in dbl], ctypes.c_int(int(kernel)),
                              ctypes.c_double(p0), ctypes.c_int(1), ctypes.c_double(p1), ctypes.c_double(tol),
                              labels.ctypes.data_as(_ip), errors.ctypes.data_as(_dp))
    if rc != 0:
        raise RuntimeError("oracle label_pairs_ss failed")
    return labels, errors


def mlp_num_params():
    l = lib()
    l.orc_mlp_num_params.restype = ctypes.c_int
    return l.orc_mlp_num_params()


def mlp_logits(params, mean, scale, x):
    params, mean, scale = (np.ascontiguousarray(a, dtype=np.float64) for a in (params, mean, scale))
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    out = np.zeros((n, 10))
    lib().orc_mlp_logits(params.ctypes.data_as(_dp), mean.ctypes.data_as(_dp),
                         scale.ctypes.data_as(_dp), ctypes.c_int(n), x.ctypes.data_as(_fp),
                         out.ctypes.data_as(_dp))
    return out


def mlp_loss_grad(params, mean, scale, x, labels, want_grad=True):
    """Batch-mean loss (and gradient) of the degree model in FP64."""
    params, mean, scale = (np.ascontiguousarray(a, dtype=np.float64) for a in (params, mean, scale))
    x = np.ascontiguousarray(x, dtype=np.float32)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    grad = np.zeros(len(params)) if want_grad else None
    l = lib()
    l.orc_mlp_loss_grad.restype = ctypes.c_double
    loss = l.orc_mlp_loss_grad(params.ctypes.data_as(_dp), mean.ctypes.data_as(_dp),
                               scale.ctypes.data_as(_dp), ctypes.c_int(x.shape[0]),
                               x.ctypes.data_as(_fp), labels.ctypes.data_as(_ip),
                               None if grad is None else grad.ctypes.data_as(_dp))
    return loss, grad


def treecode_vardeg(tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, kernel_params, singularity,
                    approximation, compute_type, theta, degree, max_source_leaf, max_target_leaf,
                    pair_degree, size_check=1.0, beta=-1.0):
    """Treecode with a degree per far-field pair (CPU FP64)."""
    tx, ty, tz, tq, sx, sy, sz, sq, sw = map(_c, (tx, ty, tz, tq, sx, sy, sz, sq, sw))
    kp = np.ascontiguousarray(np.atleast_1d(np.asarray(kernel_params, dtype=np.float64)))
    deg = np.ascontiguousarray(pair_degree, dtype=np.int32)
    out = np.zeros(len(tx))
    l = lib()
    l.orc_treecode_vardeg.restype = ctypes.c_int
    rc = l.orc_treecode_vardeg(ctypes.c_int(len(tx)), ctypes.c_int(len(sx)),
                               _p(tx), _p(ty), _p(tz), _p(tq), _p(sx), _p(sy), _p(sz), _p(sq),
                               _p(sw), _p(out), ctypes.c_int(int(kernel)), ctypes.c_int(len(kp)),
                               _p(kp), ctypes.c_int(int(singularity)),
                               ctypes.c_int(int(approximation)), ctypes.c_int(int(compute_type)),
                               ctypes.c_double(theta), ctypes.c_int(degree),
                               ctypes.c_int(max_source_leaf), ctypes.c_int(max_target_leaf),
                               ctypes.c_double(size_check), ctypes.c_double(beta),
                               ctypes.c_int(len(deg)), deg.ctypes.data_as(_ip))
    if rc != 0:
        raise RuntimeError(f"oracle treecode_vardeg returned {rc}")
    return out
