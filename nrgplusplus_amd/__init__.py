"""nrgplusplus_amd — MI355X-native Wilson-chain NRG step behind the nrgcore API.

Python front-end over the C ABI of libnrgk.so (include/nrgk.h).  The class and method names follow
the reference's C++ API (srbhp/nrgplusplus nrgcore/include/nrgcore/nrgcore.hpp):
``NrgCore.set_parameters / add_bath_site / update_internal_state / all_eigenvalue ...``.

There is no CPU fallback: every numerical call runs the HIP kernels of libnrgk.so, and importing
the library fails loudly when it is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libnrgk.so")

NRGK_MODE_MIXED = 0
NRGK_MODE_PLAIN64 = 1

_LIB = None


class KernelStat(C.Structure):
    """nrgk_kernel_stat: device time, launches and algorithmic flops/bytes of one kernel family."""
    _fields_ = [("family", C.c_char * 32), ("ms", C.c_double), ("launches", C.c_longlong),
                ("flops", C.c_double), ("bytes", C.c_double)]


class SiteDesc(C.Structure):
    _fields_ = [
        ("n_sectors", C.c_int),
        ("n_qn", C.c_int),
        ("qn", C.POINTER(C.c_int)),
        ("dims", C.POINTER(C.c_int)),
        ("energies", C.POINTER(C.c_double)),
        ("chi", C.POINTER(C.c_double)),
        ("n_ops", C.c_int),
        ("op_nblocks", C.POINTER(C.c_int)),
        ("op_bi", C.POINTER(C.c_int)),
        ("op_bj", C.POINTER(C.c_int)),
        ("op_data", C.POINTER(C.c_double)),
    ]


class Request(C.Structure):
    _fields_ = [
        ("impurity", C.POINTER(SiteDesc)),
        ("bath", C.POINTER(SiteDesc)),
        ("observable", C.POINTER(SiteDesc)),
        ("max_kept", C.c_size_t),
        ("lam", C.c_double),
        ("first_hopping", C.c_double),
        ("n_sites", C.c_int),
        ("n_bins", C.c_int),
        ("omega", C.POINTER(C.c_double)),
        ("n_omega", C.c_int),
        ("broadening", C.c_double),
        ("fail_at_site", C.c_int),
    ]


class NrgkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("nrgk status %d: %s" % (status, msg))
        self.status = status


def build():
    """Compile libnrgk.so in-tree for gfx950 (hipcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-j8", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise ImportError("libnrgk.so is not built (run nrgplusplus_amd.build() or make -C nrgplusplus_amd)")
        L = C.CDLL(LIBPATH)
        vp, d, i, sz = C.c_void_p, C.c_double, C.c_int, C.c_size_t
        dp, ip, szp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_size_t)
        pp = C.POINTER(C.c_void_p)
        sd = C.POINTER(SiteDesc)
        sig = {
            "nrgk_last_error": (C.c_char_p, []),
            "nrgk_version": (C.c_char_p, []),
            "nrgk_site_spinhalf": (i, [d, d, d, pp]),
            "nrgk_site_two_channel": (i, [pp]),
            "nrgk_site_kondo": (i, [d, d, pp]),
            "nrgk_site_fdag_chi_spinhalf": (i, [pp]),
            "nrgk_site_double_occupancy_spinhalf": (i, [pp]),
            "nrgk_site_spinless": (i, [d, pp]),
            "nrgk_site_spinless_two_lead": (i, [d, pp]),
            "nrgk_site_resonant_two_lead": (i, [d, d, d, d, i, pp]),
            "nrgk_site_spinhalf_sc": (i, [d, d, d, d, pp]),
            "nrgk_site_free": (i, [i, pp]),
            "nrgk_site_from_hamiltonian": (i, [i, i, ip, dp, i, dp, pp]),
            "nrgk_fock_basis": (i, [i, i, dp, ip]),
            "nrgk_site_duplicate_ops": (i, [vp]),
            "nrgk_site_view": (i, [vp, sd]),
            "nrgk_site_destroy": (i, [vp]),
            "nrgk_ctx_create": (i, [i, pp]),
            "nrgk_ctx_destroy": (i, [vp]),
            "nrgk_ctx_set_capacity": (i, [vp, sz]),
            "nrgk_ctx_used_bytes": (i, [vp, szp, szp]),
            "nrgk_chain_create": (i, [vp, sd, sd, sz, i, pp]),
            "nrgk_chain_destroy": (i, [vp]),
            "nrgk_chain_set_parameters": (i, [vp, sz]),
            "nrgk_chain_set_bath": (i, [vp, sd]),
            "nrgk_chain_add_bath_site": (i, [vp, dp, i, d]),
            "nrgk_chain_update_internal_state": (i, [vp]),
            "nrgk_chain_step": (i, [vp, dp, i, d]),
            "nrgk_chains_step": (i, [pp, i, dp, i, dp, pp, i, pp, i, ip]),
            "nrgk_eigh_batched": (i, [vp, i, ip, dp, dp, dp, i]),
            "nrgk_chain_counters": (i, [vp, ip, ip]),
            "nrgk_chain_set_counter": (i, [vp, i]),
            "nrgk_chain_higher_energy_discarded": (i, [vp, ip]),
            "nrgk_chain_all_eigenvalues": (i, [vp, dp, sz, szp]),
            "nrgk_chain_relative_gs": (i, [vp, dp, sz, szp]),
            "nrgk_chain_sectors": (i, [vp, ip, ip, ip]),
            "nrgk_chain_eigenvalues_q": (i, [vp, dp, sz, szp]),
            "nrgk_chain_kept_counts": (i, [vp, ip, i, ip]),
            "nrgk_chain_block_u": (i, [vp, i, dp, sz]),
            "nrgk_chain_keep_hamiltonians": (i, [vp, i]),
            "nrgk_chain_block_h": (i, [vp, i, dp, sz]),
            "nrgk_chain_fdag_blocks": (i, [vp, i, ip, ip, ip, ip, ip, dp, sz]),
            "nrgk_chain_set_state": (i, [vp, sd, i]),
            "nrgk_chain_step_stats": (i, [vp, dp, ip, ip]),
            "nrgk_chain_refine_stats": (i, [vp, ip, ip]),
            "nrgk_chain_step_counts": (i, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
            "nrgk_chain_truncation_stats": (i, [vp, dp, ip]),
            "nrgk_chain_profile": (i, [vp, i]),
            "nrgk_chain_profile_only": (i, [vp, C.c_char_p]),
            "nrgk_chain_kernel_stats": (i, [vp, vp, i, ip]),
        }
        optional = {
            "nrgk_obs_create": (i, [vp, i, ip, ip, ip, dp, pp]),
            "nrgk_obs_destroy": (i, [vp]),
            "nrgk_obs_update": (i, [vp]),
            "nrgk_obs_blocks": (i, [vp, i, ip, ip, ip, ip, ip, dp, sz]),
            "nrgk_fdm_create": (i, [vp, vp, i, d, d, pp]),
            "nrgk_fdm_destroy": (i, [vp]),
            "nrgk_fdm_record": (i, [vp]),
            "nrgk_fdm_backward": (i, [vp, d, i, i, dp, dp, dp, dp, i, ip]),
            "nrgk_fdm_energy_points": (i, [vp, dp]),
            "nrgk_fdm_signed_weights": (i, [vp, dp, i, ip]),
            "nrgk_thermo": (i, [vp, d, i, i, dp, i, dp, dp]),
            "nrgk_fdm_save": (i, [vp, C.c_char_p, C.c_char_p]),
            "nrgk_fdm_load": (i, [vp, C.c_char_p, C.c_char_p]),
            "nrgk_h5_read": (i, [C.c_char_p, C.c_char_p, dp, sz, szp]),
            "nrgk_fdm_set_system_operator": (i, [vp, vp]),
            "nrgk_thermo_expectation": (i, [vp, d, i, i, dp, i, dp]),
            "nrgk_broaden": (i, [vp, dp, dp, i, dp, i, d, dp]),
            "nrgk_tune": (i, [vp, C.c_char_p, ip, dp, i, ip]),
            "nrgk_ctx_crossover": (i, [vp, ip, ip]),
            "nrgk_server_create": (i, [vp, i, sz, C.POINTER(vp)]),
            "nrgk_server_destroy": (i, [vp]),
            "nrgk_server_submit": (i, [vp, C.POINTER(Request), C.POINTER(C.c_longlong)]),
            "nrgk_server_poll": (i, [vp, C.c_longlong, ip]),
            "nrgk_server_wait": (i, [vp, C.c_longlong, dp, dp, dp]),
            "nrgk_server_stats": (i, [vp, ip, ip, szp, ip, szp]),
            "nrgk_zdisc_chain": (i, [d, d, d, dp, dp, i, i, i, i, i, dp, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        for name, (res, args) in optional.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype, f.argtypes = res, args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise NrgkError(rc, lib().nrgk_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


# ---------------------------------------------------------------- site models
class Site:
    """A site model (impurity or bath): sectors, energies, chi and f-dagger blocks.

    Built by the library (spinhalf / two_channel / kondo / fdag_chi_spinhalf) or from arrays.
    """

    def __init__(self, handle=None, arrays=None):
        self._h = handle
        self._arrays = arrays
        self._desc = None

    @classmethod
    def spinhalf(cls, eps, U, h=0.0):
        p = C.c_void_p()
        _check(lib().nrgk_site_spinhalf(eps, U, h, C.byref(p)))
        return cls(p)

    @classmethod
    def two_channel(cls):
        p = C.c_void_p()
        _check(lib().nrgk_site_two_channel(C.byref(p)))
        return cls(p)

    @classmethod
    def kondo(cls, J, S=0.5):
        p = C.c_void_p()
        _check(lib().nrgk_site_kondo(J, S, C.byref(p)))
        return cls(p)

    @classmethod
    def fdag_chi_spinhalf(cls):
        p = C.c_void_p()
        _check(lib().nrgk_site_fdag_chi_spinhalf(C.byref(p)))
        return cls(p)

    @classmethod
    def double_occupancy_spinhalf(cls):
        """n_up n_dn of a spinhalf site (spinhalf::getDoubleOcc, spinhalf.hpp:106-114, 215)"""
        p = C.c_void_p()
        _check(lib().nrgk_site_double_occupancy_spinhalf(C.byref(p)))
        return cls(p)

    @classmethod
    def spinless(cls, eps=0.0):
        """spinless(eps) (models/spinnless.hpp): one orbital, charge only"""
        p = C.c_void_p()
        _check(lib().nrgk_site_spinless(eps, C.byref(p)))
        return cls(p)

    @classmethod
    def spinless_two_lead(cls, eps=0.0):
        """spinnlessTwoLead(eps) (models/spinlessTwoLead.hpp): two orbitals, charge only"""
        p = C.c_void_p()
        _check(lib().nrgk_site_spinless_two_lead(eps, C.byref(p)))
        return cls(p)

    @classmethod
    def resonant_two_lead(cls, eps, v_left, v_right, U, observables=False):
        """resonantTwoLead (models/resonantTwoLead.hpp:20-74); observables=True gives the same
        sectors with the operators impurityFdag and impurityNparticle"""
        p = C.c_void_p()
        _check(lib().nrgk_site_resonant_two_lead(eps, v_left, v_right, U, int(observables), C.byref(p)))
        return cls(p)

    @classmethod
    def spinhalf_sc(cls, eps=0.0, U=0.0, h=0.0, delta=0.0):
        """spinhalfSC (models/spinhalfSC.hpp:27-123) with the pairing term delta (f_up f_dn + h.c.)"""
        p = C.c_void_p()
        _check(lib().nrgk_site_spinhalf_sc(eps, U, h, delta, C.byref(p)))
        return cls(p)

    @classmethod
    def free(cls, symmetry="charge_and_spin"):
        """a free two-orbital Wilson site (freeModel.hpp), in one of the three symmetries"""
        p = C.c_void_p()
        _check(lib().nrgk_site_free(SYMMETRY[symmetry], C.byref(p)))
        return cls(p)

    @classmethod
    def from_hamiltonian(cls, qn_state, H, ops):
        """any model: per-state quantum numbers [D][n_qn], a symmetric D x D Hamiltonian that keeps
        them, and full D x D operators; sectors diagonalised and operators rotated into the eigenbasis"""
        qn = np.ascontiguousarray(qn_state, np.int32)
        if qn.ndim == 1:
            qn = qn[:, None]
        D, nq = qn.shape
        H = np.ascontiguousarray(H, np.float64)
        assert H.shape == (D, D)
        o = np.ascontiguousarray(np.array(ops, np.float64).reshape(len(ops), D, D)) if len(ops) else np.zeros(1)
        p = C.c_void_p()
        _check(lib().nrgk_site_from_hamiltonian(D, nq, _ip(qn.reshape(-1)), _dp(H.reshape(-1)), len(ops),
                                                _dp(o.reshape(-1)), C.byref(p)))
        return cls(p)

    @classmethod
    def from_arrays(cls, qn, dims, energies, chi, ops):
        """ops: list over operators of [(bi, bj, 2-D block)]"""
        qn = np.ascontiguousarray(qn, np.int32).reshape(-1)
        dims = np.ascontiguousarray(dims, np.int32)
        energies = np.ascontiguousarray(energies, np.float64)
        chi = np.ascontiguousarray(chi, np.float64)
        nbl = np.array([len(o) for o in ops], np.int32)
        bi = np.array([b[0] for o in ops for b in o], np.int32)
        bj = np.array([b[1] for o in ops for b in o], np.int32)
        data = np.concatenate([np.ascontiguousarray(b[2], np.float64).reshape(-1) for o in ops for b in o]) \
            if len(bi) else np.zeros(1)
        arrays = dict(qn=qn, dims=dims, energies=energies, chi=chi, nbl=nbl, bi=bi, bj=bj, data=data,
                      n_qn=int(len(qn) // max(1, len(dims))))
        return cls(None, arrays)

    def duplicate_ops(self):
        _check(lib().nrgk_site_duplicate_ops(self._h))
        self._desc = None
        return self

    def desc(self):
        if self._h is not None:
            d = SiteDesc()
            _check(lib().nrgk_site_view(self._h, C.byref(d)))
            return d
        a = self._arrays
        d = SiteDesc()
        d.n_sectors = len(a["dims"])
        d.n_qn = a["n_qn"]
        d.qn = _ip(a["qn"])
        d.dims = _ip(a["dims"])
        d.energies = _dp(a["energies"])
        d.chi = _dp(a["chi"])
        d.n_ops = len(a["nbl"])
        d.op_nblocks = _ip(a["nbl"])
        d.op_bi = _ip(a["bi"])
        d.op_bj = _ip(a["bj"])
        d.op_data = _dp(a["data"])
        return d

    def export(self):
        """(qn[S, n_qn], dims[S], energies (concatenated), chi[S], ops=[[(bi, bj, block)]])"""
        d = self.desc()
        S, nq = d.n_sectors, d.n_qn
        qn = np.ctypeslib.as_array(d.qn, (S * nq,)).reshape(S, nq).copy()
        dims = np.ctypeslib.as_array(d.dims, (S,)).copy()
        en = np.ctypeslib.as_array(d.energies, (int(dims.sum()),)).copy() if dims.sum() else np.zeros(0)
        chi = np.ctypeslib.as_array(d.chi, (S,)).copy()
        nbl = np.ctypeslib.as_array(d.op_nblocks, (d.n_ops,)).copy() if d.n_ops else np.zeros(0, np.int32)
        tot = int(nbl.sum())
        bi = np.ctypeslib.as_array(d.op_bi, (tot,)).copy() if tot else np.zeros(0, np.int32)
        bj = np.ctypeslib.as_array(d.op_bj, (tot,)).copy() if tot else np.zeros(0, np.int32)
        sizes = dims[bi] * dims[bj]
        data = np.ctypeslib.as_array(d.op_data, (int(sizes.sum()),)).copy() if tot else np.zeros(0)
        ops, b, off = [], 0, 0
        for p in range(d.n_ops):
            blocks = []
            for _ in range(nbl[p]):
                r, c = dims[bi[b]], dims[bj[b]]
                blocks.append((int(bi[b]), int(bj[b]), data[off: off + r * c].reshape(r, c).copy()))
                off += r * c
                b += 1
            ops.append(blocks)
        return qn, dims, en, chi, ops

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nrgk_site_destroy(self._h)
            self._h = None


# ---------------------------------------------------------------- context
SYMMETRY = {"charge_and_spin": 0, "charge_only": 1, "spin_only": 2}


def fock_basis(n_orb, symmetry="charge_and_spin"):
    """fermionBasis (models/fermionBasis.hpp:20-442): the Jordan-Wigner f^dagger_i as full
    2^n_orb x 2^n_orb matrices and the quantum numbers of each basis state"""
    D = 1 << n_orb
    sym = SYMMETRY[symmetry]
    nq = 2 if sym == 0 else 1
    f = np.zeros(n_orb * D * D)
    q = np.zeros(D * nq, np.int32)
    _check(lib().nrgk_fock_basis(n_orb, sym, _dp(f), _ip(q)))
    return list(f.reshape(n_orb, D, D)), q.reshape(D, nq)


class Context:
    """One HIP device.  Chains allocate from its capacity (HBM admission)."""

    def __init__(self, device=0):
        p = C.c_void_p()
        _check(lib().nrgk_ctx_create(device, C.byref(p)))
        self._h = p

    def set_capacity(self, nbytes):
        _check(lib().nrgk_ctx_set_capacity(self._h, nbytes))

    def tune(self, cache_path=None):
        """autotune the eigensolver crossover (nrgk_tune): (crossover, rows of (width, blocks, jacobi ms,
        householder ms))"""
        cross, nr = C.c_int(), C.c_int()
        tab = np.zeros(4 * 32)
        _check(lib().nrgk_tune(self._h, os.fsencode(cache_path) if cache_path else None, C.byref(cross), _dp(tab),
                               32, C.byref(nr)))
        return cross.value, tab[: 4 * nr.value].reshape(nr.value, 4)

    def crossover(self):
        """(crossover, tuned): the block width above which the Householder path solves a block"""
        c, t = C.c_int(), C.c_int()
        _check(lib().nrgk_ctx_crossover(self._h, C.byref(c), C.byref(t)))
        return c.value, bool(t.value)

    def used_bytes(self):
        u, c = C.c_size_t(), C.c_size_t()
        _check(lib().nrgk_ctx_used_bytes(self._h, C.byref(u), C.byref(c)))
        return u.value, c.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nrgk_ctx_destroy(self._h)
            self._h = None


_DEFAULT_CTX = None


def default_context():
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


# ---------------------------------------------------------------- nrgcore
class NrgCore:
    """nrgcore<im_type, bath_type> (nrgcore/include/nrgcore/nrgcore.hpp:25-550) on the GPU."""

    def __init__(self, impurity, bath, max_kept=1024, mode=NRGK_MODE_MIXED, ctx=None):
        self.ctx = ctx or default_context()
        self.impurity, self.bath = impurity, bath
        di, db = impurity.desc(), bath.desc()
        p = C.c_void_p()
        _check(lib().nrgk_chain_create(self.ctx._h, C.byref(di), C.byref(db), max_kept, mode, C.byref(p)))
        self._h = p
        self.n_ops = di.n_ops
        self.n_qn = di.n_qn

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nrgk_chain_destroy(self._h)
            self._h = None

    # -- the reference's methods
    def set_parameters(self, n=1024):
        _check(lib().nrgk_chain_set_parameters(self._h, n))

    def add_bath_site(self, thopping, rescale):
        t = np.ascontiguousarray(thopping, np.float64)
        _check(lib().nrgk_chain_add_bath_site(self._h, _dp(t), len(t), float(rescale)))

    def update_internal_state(self):
        _check(lib().nrgk_chain_update_internal_state(self._h))

    def step(self, thopping, rescale):
        t = np.ascontiguousarray(thopping, np.float64)
        _check(lib().nrgk_chain_step(self._h, _dp(t), len(t), float(rescale)))

    def checkHigherEnergyDiscarded(self):
        o = C.c_int()
        _check(lib().nrgk_chain_higher_energy_discarded(self._h, C.byref(o)))
        return bool(o.value)

    @property
    def nrg_iterations_cnt(self):
        c, m = C.c_int(), C.c_int()
        _check(lib().nrgk_chain_counters(self._h, C.byref(c), C.byref(m)))
        return c.value

    @nrg_iterations_cnt.setter
    def nrg_iterations_cnt(self, v):
        _check(lib().nrgk_chain_set_counter(self._h, int(v)))

    @property
    def nrg_iterations_min(self):
        c, m = C.c_int(), C.c_int()
        _check(lib().nrgk_chain_counters(self._h, C.byref(c), C.byref(m)))
        return m.value

    @property
    def all_eigenvalue(self):
        n = C.c_size_t()
        _check(lib().nrgk_chain_all_eigenvalues(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        _check(lib().nrgk_chain_all_eigenvalues(self._h, _dp(out), n.value, C.byref(n)))
        return out

    @property
    def relativeGroundStateEnergy(self):
        n = C.c_size_t()
        _check(lib().nrgk_chain_relative_gs(self._h, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value))
        _check(lib().nrgk_chain_relative_gs(self._h, _dp(out), len(out), C.byref(n)))
        return out[: n.value]

    def sectors(self):
        """(current_sysmQ as [S, n_qn], block dimensions [S])"""
        s = C.c_int()
        _check(lib().nrgk_chain_sectors(self._h, C.byref(s), None, None))
        qn = np.zeros((s.value, self.n_qn), np.int32)
        dims = np.zeros(s.value, np.int32)
        _check(lib().nrgk_chain_sectors(self._h, C.byref(s), _ip(qn), _ip(dims)))
        return qn, dims

    @property
    def current_sysmQ(self):
        return self.sectors()[0]

    @property
    def eigenvaluesQ(self):
        _, dims = self.sectors()
        n = C.c_size_t()
        out = np.zeros(int(dims.sum()))
        _check(lib().nrgk_chain_eigenvalues_q(self._h, _dp(out), len(out), C.byref(n)))
        res, off = [], 0
        for k in dims:
            res.append(out[off: off + k].copy())
            off += k
        return res

    def kept_counts(self):
        n = C.c_int()
        _check(lib().nrgk_chain_kept_counts(self._h, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value), np.int32)
        _check(lib().nrgk_chain_kept_counts(self._h, _ip(out), len(out), C.byref(n)))
        return out[: n.value]

    @property
    def eigenvaluesQ_kept_indices(self):
        return [list(range(k)) for k in self.kept_counts()]

    def current_hamiltonQ(self, s):
        """U of sector s (eigenvectors in columns)."""
        _, dims = self.sectors()
        n = int(dims[s])
        out = np.zeros((n, n))
        _check(lib().nrgk_chain_block_u(self._h, s, _dp(out), out.size))
        return out

    def keep_hamiltonians(self, on=True):
        _check(lib().nrgk_chain_keep_hamiltonians(self._h, int(on)))

    def block_h(self, s):
        _, dims = self.sectors()
        n = int(dims[s])
        out = np.zeros((n, n))
        _check(lib().nrgk_chain_block_h(self._h, s, _dp(out), out.size))
        return out

    def get_f_dag_operator(self):
        """kept x kept blocks of each Wilson-site f-dagger: list over ops of [(i, j, block)]"""
        res = []
        for p in range(self.n_ops):
            nb = C.c_int()
            _check(lib().nrgk_chain_fdag_blocks(self._h, p, C.byref(nb), None, None, None, None, None, 0))
            k = nb.value
            bi, bj, r, c = (np.zeros(max(1, k), np.int32) for _ in range(4))
            _check(lib().nrgk_chain_fdag_blocks(self._h, p, C.byref(nb), _ip(bi), _ip(bj), _ip(r), _ip(c), None, 0))
            tot = int((r[:k] * c[:k]).sum())
            data = np.zeros(max(1, tot))
            _check(lib().nrgk_chain_fdag_blocks(self._h, p, C.byref(nb), _ip(bi), _ip(bj), _ip(r), _ip(c), _dp(data),
                                                len(data)))
            blocks, off = [], 0
            for b in range(k):
                blocks.append((int(bi[b]), int(bj[b]), data[off: off + r[b] * c[b]].reshape(r[b], c[b]).copy()))
                off += r[b] * c[b]
            res.append(blocks)
        return res

    def set_state(self, state_site, cnt):
        d = state_site.desc()
        _check(lib().nrgk_chain_set_state(self._h, C.byref(d), cnt))

    def step_stats(self):
        ms = np.zeros(4)
        sw, mc = C.c_int(), C.c_int()
        _check(lib().nrgk_chain_step_stats(self._h, _dp(ms), C.byref(sw), C.byref(mc)))
        oa, unc = C.c_int(), C.c_int()
        _check(lib().nrgk_chain_refine_stats(self._h, C.byref(oa), C.byref(unc)))
        nl, ns = C.c_longlong(), C.c_longlong()
        _check(lib().nrgk_chain_step_counts(self._h, C.byref(nl), C.byref(ns)))
        cut, ties = C.c_double(), C.c_int()
        _check(lib().nrgk_chain_truncation_stats(self._h, C.byref(cut), C.byref(ties)))
        return dict(ms_assemble=ms[0], ms_eig=ms[1], ms_truncate=ms[2], ms_rotate=ms[3], jacobi_sweeps=sw.value,
                    max_cluster=mc.value, oa_steps=oa.value, oa_unconverged=unc.value, launches=nl.value,
                    host_syncs=ns.value, e_cut=cut.value, near_ties=ties.value)

    def profile(self, enable=True, only=None):
        """Start (clearing the totals) or stop per-kernel-family device timing; `only` restricts it to
        one family."""
        _check(lib().nrgk_chain_profile_only(self._h, only.encode() if only else None))
        _check(lib().nrgk_chain_profile(self._h, 1 if enable else 0))

    def kernel_stats(self):
        """{family: dict(ms, launches, flops, bytes)} accumulated while profiling was on."""
        n = C.c_int()
        _check(lib().nrgk_chain_kernel_stats(self._h, None, 0, C.byref(n)))
        arr = (KernelStat * max(1, n.value))()
        _check(lib().nrgk_chain_kernel_stats(self._h, arr, n.value, C.byref(n)))
        return {arr[k].family.decode(): dict(ms=arr[k].ms, launches=arr[k].launches, flops=arr[k].flops,
                                             bytes=arr[k].bytes) for k in range(n.value)}


def _blocks_from(getter, h, n_ops):
    res = []
    for p in range(n_ops):
        nb = C.c_int()
        _check(getter(h, p, C.byref(nb), None, None, None, None, None, 0))
        k = nb.value
        bi, bj, r, c = (np.zeros(max(1, k), np.int32) for _ in range(4))
        _check(getter(h, p, C.byref(nb), _ip(bi), _ip(bj), _ip(r), _ip(c), None, 0))
        tot = int((r[:k].astype(np.int64) * c[:k]).sum())
        data = np.zeros(max(1, tot))
        _check(getter(h, p, C.byref(nb), _ip(bi), _ip(bj), _ip(r), _ip(c), _dp(data), len(data)))
        blocks, off = [], 0
        for b in range(k):
            blocks.append((int(bi[b]), int(bj[b]), data[off: off + r[b] * c[b]].reshape(r[b], c[b]).copy()))
            off += r[b] * c[b]
        res.append(blocks)
    return res


class SystemOperator:
    """A set of system operators rotated along the chain: update_system_operator
    (nrgcore/include/nrgcore/sysOperator.hpp:35-133).  ``site`` supplies the blocks in the impurity
    sectors (e.g. ``Site.fdag_chi_spinhalf()``)."""

    def __init__(self, chain, site):
        self.chain = chain
        _, dims, _, _, ops = site.export()
        self.n_ops = len(ops)
        nbl = np.array([len(o) for o in ops], np.int32)
        bi = np.array([b[0] for o in ops for b in o], np.int32)
        bj = np.array([b[1] for o in ops for b in o], np.int32)
        data = np.concatenate([b[2].reshape(-1) for o in ops for b in o])
        p = C.c_void_p()
        _check(lib().nrgk_obs_create(chain._h, self.n_ops, _ip(nbl), _ip(bi), _ip(bj), _dp(data), C.byref(p)))
        self._h = p

    def update(self):
        _check(lib().nrgk_obs_update(self._h))

    def blocks(self):
        return _blocks_from(lib().nrgk_obs_blocks, self._h, self.n_ops)

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nrgk_obs_destroy(self._h)
            self._h = None


def update_system_operator(chain, op):
    """update_system_operator(nrg_object, operators) (sysOperator.hpp:35-37)"""
    op.update()


class FdmRecorder:
    """Device-resident hand-off of the forward pass (what NrgData saves per iteration) plus the T = 0
    FDM backward pass (fdmSpectrum) and the thermodynamics (fdmThermodynamics)."""

    def __init__(self, chain, op=None, n_bins=100000, e_min=1e-10, e_max=10.0):
        self.chain, self.op = chain, op
        self.sysop = None
        self.n_bins = n_bins
        p = C.c_void_p()
        _check(lib().nrgk_fdm_create(chain._h, op._h if op is not None else None, n_bins, e_min, e_max, C.byref(p)))
        self._h = p

    def record(self):
        _check(lib().nrgk_fdm_record(self._h))

    def energy_points(self):
        out = np.zeros(self.n_bins)
        _check(lib().nrgk_fdm_energy_points(self._h, _dp(out)))
        return out

    def backward(self, lam, scale_offset=-1, first_cnt=0, max_shells=4096):
        n_ops = self.op.n_ops
        pos = np.zeros(n_ops * self.n_bins)
        neg = np.zeros(n_ops * self.n_bins)
        tr = np.zeros(max_shells)
        sm = np.zeros(max_shells)
        ns = C.c_int()
        _check(lib().nrgk_fdm_backward(self._h, lam, scale_offset, first_cnt, _dp(pos), _dp(neg), _dp(tr), _dp(sm),
                                       max_shells, C.byref(ns)))
        return pos.reshape(n_ops, self.n_bins), neg.reshape(n_ops, self.n_bins), tr[: ns.value], sm[: ns.value]

    def signed_weights(self):
        """per operator, the signed weights of the last backward pass summed over its shells (the
        complete-basis sum rule: <{B, B+}>)"""
        n = C.c_int(0)
        _check(lib().nrgk_fdm_signed_weights(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        _check(lib().nrgk_fdm_signed_weights(self._h, _dp(out), n.value, C.byref(n)))
        return out

    def thermodynamics(self, lam, T, scale_offset=-1, first_cnt=0):
        T = np.ascontiguousarray(T, np.float64)
        S = np.zeros(len(T))
        Cv = np.zeros(len(T))
        _check(lib().nrgk_thermo(self._h, lam, scale_offset, first_cnt, _dp(T), len(T), _dp(S), _dp(Cv)))
        return S, Cv

    def save(self, path, op_group="/nparticle"):
        """NrgData::saveCurrentData + saveqOperator of every recorded shell (nrgData.hpp:134-205), in
        the reference's HDF5 layout; op_group=None leaves the observable out"""
        _check(lib().nrgk_fdm_save(self._h, os.fsencode(path), op_group.encode() if op_group else None))

    def load(self, path, op_group="/nparticle"):
        """NrgData::loadCurrentData + loadqOperator (nrgData.hpp:214-293) of every shell of a file save()
        wrote: the recorder's shells are replaced by the file's"""
        _check(lib().nrgk_fdm_load(self._h, os.fsencode(path), op_group.encode() if op_group else None))

    def setSystemOperator(self, op):
        """fdmThermodynamics::setSystemOperator (static/fdmThermodynamics.hpp:39-41): shells recorded
        from now on keep the diagonals of ``op`` (a SystemOperator updated before each record)."""
        _check(lib().nrgk_fdm_set_system_operator(self._h, op._h))
        self.sysop = op

    def expectation_values(self, lam, T, scale_offset=-1, first_cnt=0):
        """qsysOPeratorExpValue (fdmThermodynamics.hpp:191-211): [n_ops][len(T)]"""
        if self.sysop is None:
            raise RuntimeError("expectation_values: no system operator set (setSystemOperator)")
        T = np.ascontiguousarray(T, np.float64)
        out = np.zeros(self.sysop.n_ops * len(T))
        _check(lib().nrgk_thermo_expectation(self._h, lam, scale_offset, first_cnt, _dp(T), len(T), _dp(out)))
        return out.reshape(self.sysop.n_ops, len(T))

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nrgk_fdm_destroy(self._h)
            self._h = None


def chains_step(chains, t, rescale, observables=None, recorders=None, n_threads=0):
    """One Wilson site on each of several independent chains, run concurrently (a host worker and
    the chain's own stream per chain).  Chain i does what the reference's FDM drivers do per site
    (examples/fdmSpectrumSiam/main.cpp:86-95): add_bath_site(t[i], rescale[i]), update of its
    observables (observables[i], a list), record into recorders[i] when higher energies were
    discarded, update_internal_state.  Returns a bool array: which chains recorded a shell."""
    n = len(chains)
    if n == 0:
        return np.zeros(0, bool)
    t = np.ascontiguousarray(np.asarray(t, np.float64).reshape(n, -1))
    r = np.ascontiguousarray(np.broadcast_to(np.asarray(rescale, np.float64), (n,)))
    obs = observables if observables is not None else [[] for _ in range(n)]
    n_obs = len(obs[0])
    if len(obs) != n or any(len(o) != n_obs for o in obs):
        raise ValueError("chains_step: one list of equally many observables per chain")
    ch = (C.c_void_p * n)(*[c._h.value for c in chains])
    ob = (C.c_void_p * max(1, n * n_obs))(*[o._h.value for row in obs for o in row])
    rec = None
    if recorders is not None:
        if len(recorders) != n:
            raise ValueError("chains_step: one recorder (or None) per chain")
        rec = (C.c_void_p * n)(*[(x._h.value if x is not None else None) for x in recorders])
    out = np.zeros(n, np.int32)
    _check(lib().nrgk_chains_step(ch, n, _dp(t), t.shape[1], _dp(r), ob, n_obs, rec, int(n_threads), _ip(out)))
    return out.astype(bool)


def eigh_batched(blocks, mode=None, ctx=None):
    """qmatrix::diag (nrgcore/include/utils/qmatrix.hpp:375-404) for a list of symmetric blocks in one
    call, through the chains' batched eigensolver: a list of (w, U), w ascending, the eigenvectors
    in the columns of U."""
    ctx = ctx or default_context()
    mode = NRGK_MODE_MIXED if mode is None else mode
    mats = [np.ascontiguousarray(b, np.float64) for b in blocks]
    for m in mats:
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise ValueError("eigh_batched: every block must be square")
    dims = np.array([m.shape[0] for m in mats], np.int32)
    H = np.concatenate([m.reshape(-1) for m in mats]) if mats else np.zeros(0)
    w = np.zeros(max(1, int(dims.sum())))
    U = np.zeros(max(1, H.size))
    _check(lib().nrgk_eigh_batched(ctx._h, len(dims), _ip(dims), _dp(H if H.size else np.zeros(1)), _dp(w), _dp(U),
                                   int(mode)))
    out, eo, uo = [], 0, 0
    for n in dims:
        n = int(n)
        out.append((w[eo:eo + n].copy(), U[uo:uo + n * n].reshape(n, n).copy()))
        eo += n
        uo += n * n
    return out


def broaden(x, w, omega, b, ctx=None):
    """log-Gaussian broadening (examples/fdmSpectrumSiam/spec.py:27-44), on the device."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(x, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    omega = np.ascontiguousarray(omega, np.float64)
    out = np.zeros(len(omega))
    _check(lib().nrgk_broaden(ctx._h, _dp(x), _dp(w), len(x), _dp(omega), len(omega), float(b), _dp(out)))
    return out


def h5_read(path, dataset):
    """a one-dimensional dataset of an HDF5 file, as doubles"""
    n = C.c_size_t()
    _check(lib().nrgk_h5_read(os.fsencode(path), dataset.encode(), None, 0, C.byref(n)))
    out = np.zeros(max(1, n.value))
    _check(lib().nrgk_h5_read(os.fsencode(path), dataset.encode(), _dp(out), len(out), C.byref(n)))
    return out[: n.value]


def z_chain(z, lam, n_sites, D=1.0, dos=None, quad="exact", n_intervals=0, normalize=True):
    """Wilson chain (t_n Lambda^{n/2}, eps_n Lambda^{n/2}) of the z-shifted logarithmic discretisation
    of a bath (independFiles/zAverage/hoppingElement.cpp:16-105, zAverage.hpp:12-195).  dos: None for
    the flat band, or (x, rho) tabulated on ascending x over [-D, D]; quad: "exact" or "reference"
    (the reference's 100-point left sums)."""
    t = np.zeros(n_sites)
    e = np.zeros(n_sites)
    if dos is None:
        x = r = np.zeros(1)
        nx = 0
    else:
        x = np.ascontiguousarray(dos[0], np.float64)
        r = np.ascontiguousarray(dos[1], np.float64)
        nx = len(x)
    q = {"exact": 0, "reference": 1}[quad]
    _check(lib().nrgk_zdisc_chain(z, lam, D, _dp(x), _dp(r), nx, q, n_intervals, int(normalize), n_sites, _dp(t),
                                  _dp(e)))
    return t, e


def hopping(n, lam):
    """Wilson-chain hopping (examples/rgflowSIAM/main.cpp:5-10)."""
    return 0.5 * (1.0 + 1.0 / lam) * (1.0 - lam ** (-n - 1)) / np.sqrt(
        (1.0 - lam ** (-2.0 * n - 1)) * (1.0 - lam ** (-2.0 * n - 3)))


# ---------------------------------------------------------------- request loop
class Server:
    """Spectral-function requests served concurrently (nrgk_server_*): admission by device-memory
    footprint and slots, one chain and stream per running request, first come first served."""

    def __init__(self, n_slots=8, hbm_budget=0, ctx=None):
        self.ctx = ctx or default_context()
        p = C.c_void_p()
        _check(lib().nrgk_server_create(self.ctx._h, n_slots, hbm_budget, C.byref(p)))
        self._h = p
        self._n_omega = {}

    def submit(self, impurity, bath, observable, max_kept, lam, first_hopping, n_sites, n_bins, omega,
               broadening=0.6, fail_at_site=-1):
        """queues one request; returns its id.  impurity, bath: Site; observable: Site whose operators are
        the observable blocks (e.g. Site.fdag_chi_spinhalf())"""
        om = np.ascontiguousarray(omega, dtype=np.float64)
        di, db, do = impurity.desc(), bath.desc(), observable.desc()
        r = Request(C.pointer(di), C.pointer(db), C.pointer(do), max_kept, lam, first_hopping, n_sites, n_bins,
                    _dp(om), len(om), broadening, fail_at_site)
        rid = C.c_longlong()
        _check(lib().nrgk_server_submit(self._h, C.byref(r), C.byref(rid)))
        self._n_omega[rid.value] = len(om)
        return rid.value

    def done(self, rid):
        d = C.c_int()
        _check(lib().nrgk_server_poll(self._h, rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid):
        """(A_pos, A_neg, info) of request rid: info = dict(latency_ms, queue_ms, total_weight, iterations,
        shells, bytes); raises NrgkError if the request failed"""
        n = self._n_omega.pop(rid)
        ap, an, info = np.zeros(n), np.zeros(n), np.zeros(6)
        _check(lib().nrgk_server_wait(self._h, rid, _dp(ap), _dp(an), _dp(info)))
        return ap, an, dict(latency_ms=info[0], queue_ms=info[1], total_weight=info[2], iterations=int(info[3]),
                            shells=int(info[4]), bytes=int(info[5]))

    def stats(self):
        r, q, pk = C.c_int(), C.c_int(), C.c_int()
        res, pres = C.c_size_t(), C.c_size_t()
        _check(lib().nrgk_server_stats(self._h, C.byref(r), C.byref(q), C.byref(res), C.byref(pk), C.byref(pres)))
        return dict(running=r.value, queued=q.value, reserved=res.value, peak_running=pk.value,
                    peak_reserved=pres.value)

    def close(self):
        if getattr(self, "_h", None) is not None:
            _check(lib().nrgk_server_destroy(self._h))
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.nrgk_server_destroy(self._h)
            self._h = None
