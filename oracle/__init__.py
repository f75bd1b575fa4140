"""ctypes binding of the FP64 oracle (liboracle.so).

Test infrastructure: only tests/, the smoke check and the bench's CPU-baseline leg import this
package.  The product path never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcDriver(C.Structure):
    _fields_ = [
        ("imp_kind", C.c_int),
        ("eps", C.c_double),
        ("U", C.c_double),
        ("h", C.c_double),
        ("J", C.c_double),
        ("S", C.c_double),
        ("bath_kind", C.c_int),
        ("dup_ops", C.c_int),
        ("Lambda", C.c_double),
        ("first_V", C.c_double),
        ("nsites", C.c_int),
        ("K", C.c_long),
    ]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, d, i, l = C.c_void_p, C.c_double, C.c_int, C.c_long
        dp, ip, lp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_lapack_library": (C.c_char_p, []),
            "orc_set_threads": (None, [i]),
            "orc_model_spinhalf": (vp, [d, d, d]),
            "orc_model_two_channel": (vp, []),
            "orc_model_kondo": (vp, [d, d]),
            "orc_model_dup_ops": (None, [vp]),
            "orc_model_free": (None, [vp]),
            "orc_model_n_sectors": (i, [vp]),
            "orc_model_n_ops": (i, [vp]),
            "orc_model_sectors": (None, [vp, ip, ip, dp, dp]),
            "orc_model_op_nblocks": (i, [vp, i]),
            "orc_model_op_blocks": (None, [vp, i, ip, ip, dp]),
            "orc_chain_new": (vp, [vp, vp, l]),
            "orc_chain_free": (None, [vp]),
            "orc_chain_keep_h": (None, [vp, i]),
            "orc_add_bath_site": (i, [vp, dp, i, d]),
            "orc_update_internal_state": (i, [vp]),
            "orc_cnt": (i, [vp]),
            "orc_syevd_failures": (i, [vp]),
            "orc_step_times": (None, [vp, dp]),
            "orc_n_all_eigenvalues": (l, [vp]),
            "orc_all_eigenvalues": (None, [vp, dp]),
            "orc_n_relgs": (l, [vp]),
            "orc_relgs": (None, [vp, dp]),
            "orc_n_sectors": (i, [vp]),
            "orc_sector_dims": (None, [vp, ip]),
            "orc_eigenvalues_q": (None, [vp, dp]),
            "orc_n_kept_sectors": (i, [vp]),
            "orc_kept_counts": (None, [vp, ip]),
            "orc_n_qn": (i, [vp]),
            "orc_n_pre_sectors": (i, [vp]),
            "orc_pre_qn": (None, [vp, ip]),
            "orc_n_cur_sectors": (i, [vp]),
            "orc_cur_qn": (None, [vp, ip]),
            "orc_block_h": (None, [vp, i, dp]),
            "orc_block_u": (None, [vp, i, dp]),
            "orc_n_ops": (i, [vp]),
            "orc_fdag_nblocks": (i, [vp, i]),
            "orc_fdag_block_info": (None, [vp, i, ip, ip, ip, ip]),
            "orc_fdag_block": (None, [vp, i, i, dp]),
            "orc_run_rgflow": (i, [C.POINTER(OrcDriver), dp, l, lp]),
            "orc_run_thermo": (i, [C.POINTER(OrcDriver), i, dp, i, dp, dp, i]),
            "orc_run_thermo_t": (i, [C.POINTER(OrcDriver), dp, i, dp, i, dp, dp]),
            "orc_run_thermo_ops": (i, [C.POINTER(OrcDriver), i, dp, i, i, dp, dp, dp, ip]),
            "orc_run_fdm": (i, [C.POINTER(OrcDriver), i, dp, dp, dp, dp, ip]),
            "orc_run_fdm_signed": (i, [C.POINTER(OrcDriver), i, dp, dp, dp, dp, ip, dp]),
            "orc_set_fdm_plant": (None, [i]),
            "orc_broaden": (None, [dp, dp, i, dp, i, d, dp]),
            "orc_time_forward": (i, [C.POINTER(OrcDriver), i, i, dp]),
            "orc_time_rgflow": (i, [C.POINTER(OrcDriver), i, i, dp]),
            "orc_time_forward2": (i, [C.POINTER(OrcDriver), i, i, i, dp, dp]),
            "orc_hopping": (d, [i, d]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def hopping(n, L):
    return lib().orc_hopping(n, L)


class Model:
    """A site model (impurity or bath) built by the oracle."""

    def __init__(self, kind="spinhalf", eps=0.0, U=0.0, h=0.0, J=0.0, S=0.5, dup_ops=False):
        L = lib()
        if kind == "spinhalf":
            self.h = L.orc_model_spinhalf(eps, U, h)
        elif kind == "two_channel":
            self.h = L.orc_model_two_channel()
        elif kind == "kondo":
            self.h = L.orc_model_kondo(J, S)
        else:
            raise ValueError(kind)
        if not self.h:
            _check(-1)
        if dup_ops:
            L.orc_model_dup_ops(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_model_free(self.h)
            self.h = None

    def export(self):
        """(qn[S,2], dims[S], energies(concat), chi[S], ops=[list of (bi,bj,mat)])"""
        L = lib()
        S = L.orc_model_n_sectors(self.h)
        qn = np.zeros((S, 2), np.int32)
        dims = np.zeros(S, np.int32)
        # energies length = sum dims; first query dims with a big buffer
        en = np.zeros(4096)
        chi = np.zeros(S)
        L.orc_model_sectors(self.h, _ip(qn), _ip(dims), _dp(en), _dp(chi))
        en = en[: dims.sum()].copy()
        ops = []
        for p in range(L.orc_model_n_ops(self.h)):
            nb = L.orc_model_op_nblocks(self.h, p)
            bi = np.zeros(nb, np.int32)
            bj = np.zeros(nb, np.int32)
            data = np.zeros(65536)
            L.orc_model_op_blocks(self.h, p, _ip(bi), _ip(bj), _dp(data))
            blocks, off = [], 0
            for b in range(nb):
                r, c = dims[bi[b]], dims[bj[b]]
                blocks.append((int(bi[b]), int(bj[b]), data[off: off + r * c].reshape(r, c).copy()))
                off += r * c
            ops.append(blocks)
        return qn, dims, en, chi, ops


class Chain:
    """Step-by-step FP64 oracle chain (restates nrgcore<imp,bath>)."""

    def __init__(self, imp, bath, K, keep_h=False):
        self.imp, self.bath = imp, bath  # keep alive
        self.h = lib().orc_chain_new(imp.h, bath.h, K)
        if not self.h:
            _check(-1)
        if keep_h:
            lib().orc_chain_keep_h(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_chain_free(self.h)
            self.h = None

    def add_bath_site(self, t, rescale):
        t = np.ascontiguousarray(t, dtype=np.float64)
        _check(lib().orc_add_bath_site(self.h, _dp(t), len(t), rescale))

    def update_internal_state(self):
        _check(lib().orc_update_internal_state(self.h))

    @property
    def cnt(self):
        return lib().orc_cnt(self.h)

    def all_eigenvalue(self):
        n = lib().orc_n_all_eigenvalues(self.h)
        out = np.zeros(n)
        lib().orc_all_eigenvalues(self.h, _dp(out))
        return out

    def relative_gs(self):
        n = lib().orc_n_relgs(self.h)
        out = np.zeros(n)
        lib().orc_relgs(self.h, _dp(out))
        return out

    def sector_dims(self):
        S = lib().orc_n_sectors(self.h)
        d = np.zeros(S, np.int32)
        lib().orc_sector_dims(self.h, _ip(d))
        return d

    def eigenvalues_q(self):
        d = self.sector_dims()
        out = np.zeros(int(d.sum()))
        lib().orc_eigenvalues_q(self.h, _dp(out))
        res, off = [], 0
        for n in d:
            res.append(out[off: off + n].copy())
            off += n
        return res

    def kept_counts(self):
        S = lib().orc_n_kept_sectors(self.h)
        k = np.zeros(S, np.int32)
        lib().orc_kept_counts(self.h, _ip(k))
        return k

    def pre_qn(self):
        S = lib().orc_n_pre_sectors(self.h)
        q = np.zeros((S, 2), np.int32)
        lib().orc_pre_qn(self.h, _ip(q))
        return q

    def cur_qn(self):
        S = lib().orc_n_cur_sectors(self.h)
        q = np.zeros((S, 2), np.int32)
        lib().orc_cur_qn(self.h, _ip(q))
        return q

    def block_h(self, s, n):
        out = np.zeros((n, n))
        lib().orc_block_h(self.h, s, _dp(out))
        return out

    def block_u(self, s, n):
        out = np.zeros((n, n))
        lib().orc_block_u(self.h, s, _dp(out))
        return out

    def fdag(self):
        """list over ops of [(i, j, full block)]"""
        L = lib()
        ops = []
        for p in range(L.orc_n_ops(self.h)):
            nb = L.orc_fdag_nblocks(self.h, p)
            bi, bj, r, c = (np.zeros(nb, np.int32) for _ in range(4))
            L.orc_fdag_block_info(self.h, p, _ip(bi), _ip(bj), _ip(r), _ip(c))
            blocks = []
            for b in range(nb):
                m = np.zeros((r[b], c[b]))
                L.orc_fdag_block(self.h, p, b, _dp(m))
                blocks.append((int(bi[b]), int(bj[b]), m))
            ops.append(blocks)
        return ops

    def step_times(self):
        out = np.zeros(4)
        lib().orc_step_times(self.h, _dp(out))
        return out


def driver(imp_kind=0, eps=0.0, U=0.0, h=0.0, J=0.0, S=0.5, bath_kind=0, dup_ops=0, Lambda=2.0,
           first_V=0.0, nsites=10, K=64):
    return OrcDriver(imp_kind, eps, U, h, J, S, bath_kind, dup_ops, Lambda, first_V, nsites, K)


def run_rgflow(drv, cap=None):
    cap = cap or drv.nsites * 40 * (drv.K + 64)
    out = np.zeros(cap)
    counts = np.zeros(drv.nsites, np.int64)
    _check(lib().orc_run_rgflow(C.byref(drv), _dp(out), cap, counts.ctypes.data_as(C.POINTER(C.c_long))))
    res, off = [], 0
    for n in counts:
        res.append(out[off: off + n].copy())
        off += n
    return res


def run_thermo(drv, T, min_iter0, low_clamp_zero=False):
    T = np.ascontiguousarray(T, dtype=np.float64)
    S = np.zeros(len(T))
    Cv = np.zeros(len(T))
    _check(lib().orc_run_thermo(C.byref(drv), min_iter0, _dp(T), len(T), _dp(S), _dp(Cv), int(low_clamp_zero)))
    return S, Cv


def run_thermo_hoppings(drv, hop, T, min_iter0):
    """run_thermo with the chain's hopping per site given (e.g. a z-shifted discretisation)"""
    T = np.ascontiguousarray(T, dtype=np.float64)
    hop = np.ascontiguousarray(hop, dtype=np.float64)
    assert len(hop) >= drv.nsites
    S = np.zeros(len(T))
    Cv = np.zeros(len(T))
    _check(lib().orc_run_thermo_t(C.byref(drv), _dp(hop), min_iter0, _dp(T), len(T), _dp(S), _dp(Cv)))
    return S, Cv


def run_thermo_ops(drv, T, min_iter0, op_kind):
    """entropy, specific heat and the thermal expectation values [n_ops][len(T)] of the system
    operators op_kind 0: {n_up n_dn}, 1: {n_up, n_dn, identity} of a spinhalf impurity"""
    T = np.ascontiguousarray(T, dtype=np.float64)
    S = np.zeros(len(T))
    Cv = np.zeros(len(T))
    V = np.zeros(3 * len(T))
    nq = C.c_int()
    _check(lib().orc_run_thermo_ops(C.byref(drv), min_iter0, _dp(T), len(T), op_kind, _dp(S), _dp(Cv), _dp(V),
                                    C.byref(nq)))
    return S, Cv, V[: nq.value * len(T)].reshape(nq.value, len(T))


def run_fdm(drv, nbins, signed=False):
    """(pos, neg, traces, sums) per operator / shell; with signed=True also the per-operator signed
    weights summed over all shells (before |.|, the 1e-20 cut and the bin range)"""
    nops = 2
    pos = np.zeros(nops * nbins)
    neg = np.zeros(nops * nbins)
    tr = np.zeros(drv.nsites + 2)
    sums = np.zeros(drv.nsites + 2)
    sg = np.zeros(nops)
    ns = C.c_int(0)
    _check(lib().orc_run_fdm_signed(C.byref(drv), nbins, _dp(pos), _dp(neg), _dp(tr), _dp(sums), C.byref(ns),
                                    _dp(sg)))
    n = ns.value
    out = (pos.reshape(nops, nbins), neg.reshape(nops, nbins), tr[:n], sums[:n])
    return out + (sg,) if signed else out


def set_fdm_plant(k):
    """plant a mistake in the oracle's FDM backward pass (tests of the tests only): 0 none, 1 rho
    rotated as U^T rho U, 2 the last bath state dropped from the partial trace"""
    lib().orc_set_fdm_plant(int(k))


def time_rgflow(drv, n_skip, n_time):
    """Seconds spent in loop sites [n_skip, n_skip + n_time) of the RG-flow driver form (first site
    {V..V}, then add_bath_site + update_internal_state per site)."""
    s = C.c_double(0)
    _check(lib().orc_time_rgflow(C.byref(drv), n_skip, n_time, C.byref(s)))
    return s.value


def time_forward(drv, n_skip, n_time):
    """Seconds spent in loop sites [n_skip, n_skip + n_time) of the faithful FDM forward pass."""
    s = C.c_double(0)
    _check(lib().orc_time_forward(C.byref(drv), n_skip, n_time, C.byref(s)))
    return s.value


def time_forward_with_single_core(drv, n_skip, n_time, n_time1):
    """time_forward of sites [n_skip, n_skip + n_time), then the next n_time1 sites of the same chain
    at one thread: (seconds, seconds_one_thread).  The caller restores its thread count."""
    s, s1 = C.c_double(0), C.c_double(0)
    _check(lib().orc_time_forward2(C.byref(drv), n_skip, n_time, n_time1, C.byref(s), C.byref(s1)))
    return s.value, s1.value


def set_threads(n):
    lib().orc_set_threads(int(n))


def broaden(x, w, omega, b):
    x = np.ascontiguousarray(x, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    omega = np.ascontiguousarray(omega, np.float64)
    out = np.zeros(len(omega))
    lib().orc_broaden(_dp(x), _dp(w), len(x), _dp(omega), len(omega), b, _dp(out))
    return out


def temperature_grid(vmin, vmax, n):
    d = (n - 1.0) / np.log(vmax / vmin)
    return vmin * np.exp(np.abs(np.arange(n)) * 1.0 / d)
