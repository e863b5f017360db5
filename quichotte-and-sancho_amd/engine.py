"""ctypes binding of libqs_hip.so (include/qs_hip.h).

ScfEngine mirrors the reference's SCF module interface for one bias point:
  solve_scf(V, fresh)       src/scf.f90:78-151
  current()                 src/observables.f90:22-39 (Current)
  g_less_obs(), dosd()      the final-sweep observables (src/scf.f90:134-140)
  init2(V), residual2(V,x)  the order-2 seams (src/scf.f90:116-122,158-188)
  sigma()                   Sigma1 and Sigma2 of the last residual

SanchoSelfEnergies mirrors SANCHO's SelfEnergies module: build_self_energies()
(SANCHO/src/self_energies.f90:38-46) after complete_input_gf
(SANCHO/src/green_functions.f90:163-173), for orders 2 and 3.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "lib", "libqs_hip.so")

QS_ERRORS = {1: "singular matrix", 2: "bad argument", 3: "HIP error", 4: "out of device memory",
             5: "unsupported size", 6: "bad state"}


class QsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("qs error %d (%s): %s" % (code, QS_ERRORS.get(code, "?"), msg))
        self.code = code


class _Problem(ctypes.Structure):
    _fields_ = [
        ("norb", ctypes.c_int32), ("n_omega", ctypes.c_int32),
        ("w_init", ctypes.c_double), ("delta", ctypes.c_double),
        ("H", ctypes.c_void_p), ("gammaL", ctypes.c_void_p), ("gammaR", ctypes.c_void_p),
        ("U_site", ctypes.c_void_p), ("order", ctypes.c_int32),
        ("beta", ctypes.c_double), ("mu", ctypes.c_double),
        ("n_atoms", ctypes.c_int32), ("atom_norb", ctypes.c_void_p),
        ("alpha_mix", ctypes.c_double), ("nhist", ctypes.c_int32), ("maxscf", ctypes.c_int32),
        ("tolf", ctypes.c_double), ("tolf_order1", ctypes.c_double),
        ("single_history", ctypes.c_int32), ("device", ctypes.c_int32),
    ]


class _Result(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("n_iter", ctypes.c_int32), ("fmax", ctypes.c_double),
                ("runaway", ctypes.c_int32), ("unphysical", ctypes.c_int32), ("i_worst", ctypes.c_int32),
                ("n_worst", ctypes.c_double), ("w_worst", ctypes.c_double)]


_lib = None


def load_library(path=SO):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libqs_hip.so is not built (%s); run __graft_entry__.build()" % path)
    L = ctypes.CDLL(path)
    P, D, I32, I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int64
    L.qs_create.argtypes = [P, ctypes.POINTER(P)]
    L.qs_destroy.argtypes = [P]
    L.qs_last_error.argtypes = [P]
    L.qs_last_error.restype = ctypes.c_char_p
    L.qs_sizes.argtypes = [P] + [P] * 6
    L.qs_device_bytes.argtypes = [P]
    L.qs_device_bytes.restype = I64
    L.qs_solve_scf.argtypes = [P, D, I32, P, P, P]
    L.qs_current.argtypes = [P, P, P]
    L.qs_g_less_obs.argtypes = [P, P]
    L.qs_dosd.argtypes = [P, P]
    L.qs_trUD.argtypes = [P, P, P]
    L.qs_init2.argtypes = [P, D, P]
    L.qs_residual2.argtypes = [P, D, I64, P, P, P]
    L.qs_get_sigma.argtypes = [P, P, P, P]
    L.qs_final_sweep.argtypes = [P, D, P]
    L.qs_footprint_bytes.argtypes = [P]
    L.qs_footprint_bytes.restype = I64
    L.qs_device_memory.argtypes = [I32, P, P]
    L.qs_profile.argtypes = [P, I32]
    L.qs_profile_read.argtypes = [P, P, P, I32]
    L.qs_dyson_choice.argtypes = [P, P, P, I32]
    L.qs_reset_problem.argtypes = [P, P]
    L.qs_set_step_mode.argtypes = [P, I32]
    L.qs_step_info.argtypes = [P, P, P]
    L.qs_set_precision.argtypes = [P, I32, D]
    L.qs_precision_steps.argtypes = [P, P, P]
    L.qs_dyson_shape.argtypes = [P, P]
    L.qs_dyson_shape.restype = ctypes.c_char_p
    L.qs_stage_name.argtypes = [I32]
    L.qs_stage_name.restype = ctypes.c_char_p
    L.qs_sancho_create.argtypes = [I32, I32, I32, I32, ctypes.POINTER(P)]
    L.qs_sancho_destroy.argtypes = [P]
    L.qs_sancho_last_error.argtypes = [P]
    L.qs_sancho_last_error.restype = ctypes.c_char_p
    L.qs_sancho_device_bytes.argtypes = [P]
    L.qs_sancho_device_bytes.restype = I64
    L.qs_sancho_self_energies.argtypes = [P, D, P, P, P, P, P, P, P, P, P, I32]
    L.qs_sancho_static_bytes.argtypes = []
    L.qs_sancho_static_bytes.restype = I64
    L.qs_sancho_static_need.argtypes = [I32, I32, I32]
    L.qs_sancho_static_need.restype = I64
    L.qs_sancho_self_energies_static.argtypes = [I32, I32, I32, D, P, P, P, P, P, P, P, P, P, P]
    L.qs_sancho_profile.argtypes = [P, I32]
    L.qs_sancho_profile_read.argtypes = [P, P, P]
    L.qs_group_create.argtypes = [I32, ctypes.POINTER(P)]
    L.qs_group_destroy.argtypes = [P]
    L.qs_group_last_error.argtypes = [P]
    L.qs_group_last_error.restype = ctypes.c_char_p
    L.qs_group_add.argtypes = [P, P, D, I32]
    L.qs_group_members.argtypes = [P]
    L.qs_group_step.argtypes = [P, I32, P, P, P, P]
    L.qs_group_step_begin.argtypes = [P]
    L.qs_group_step_end.argtypes = [P, I32, P, P, P, P]
    L.qs_group_info.argtypes = [P, P, P, P, P]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class ScfProblem:
    """The inputs of solve_scf for one system (host arrays, reference conventions)."""
    H: np.ndarray                 # (norb, norb) complex
    gammaL: np.ndarray
    gammaR: np.ndarray
    U_site: np.ndarray            # (norb,), U on the spin-up index of each site
    order: int
    n_omega: int
    w_init: float
    delta: float
    beta: float
    mu: float
    alpha_mix: float = 0.5
    nhist: int = 16
    maxscf: int = 300
    tolf: float = 1e-6
    tolf_order1: float = 1e-8
    single_history: bool = True
    atom_norb: np.ndarray = None
    device: int = 0


@dataclass
class ScfResult:
    converged: bool
    n_iter: int
    fmax: float
    runaway: bool
    unphysical: bool
    i_worst: int
    n_worst: float
    w_worst: float
    log_fmax: np.ndarray = field(default=None, repr=False)
    log_rms: np.ndarray = field(default=None, repr=False)


def _make_problem(prob):
    """(_Problem, arrays that must stay alive while it is used)"""
    H = np.asfortranarray(prob.H, dtype=complex)
    GL = np.asfortranarray(prob.gammaL, dtype=complex)
    GR = np.asfortranarray(prob.gammaR, dtype=complex)
    U = np.ascontiguousarray(prob.U_site, dtype=float)
    an = None if prob.atom_norb is None else np.ascontiguousarray(prob.atom_norb, dtype=np.int32)
    p = _Problem()
    p.norb = H.shape[0]
    p.n_omega = prob.n_omega
    p.w_init = prob.w_init
    p.delta = prob.delta
    p.H = _ptr(H)
    p.gammaL = _ptr(GL)
    p.gammaR = _ptr(GR)
    p.U_site = _ptr(U)
    p.order = prob.order
    p.beta = prob.beta
    p.mu = prob.mu
    p.n_atoms = 0 if an is None else len(an)
    p.atom_norb = _ptr(an)
    p.alpha_mix = prob.alpha_mix
    p.nhist = prob.nhist
    p.maxscf = prob.maxscf
    p.tolf = prob.tolf
    p.tolf_order1 = prob.tolf_order1
    p.single_history = int(bool(prob.single_history))
    p.device = prob.device
    return p, (H, GL, GR, U, an)


def footprint_bytes(prob: ScfProblem):
    """Device bytes a solve of `prob` needs (qs_footprint_bytes)."""
    L = load_library()
    p, keep = _make_problem(prob)
    b = L.qs_footprint_bytes(ctypes.byref(p))
    if b < 0:
        raise QsError(2, "invalid problem")
    return b


def device_memory(device=0):
    L = load_library()
    f, t = ctypes.c_int64(), ctypes.c_int64()
    rc = L.qs_device_memory(device, ctypes.byref(f), ctypes.byref(t))
    if rc != 0:
        raise QsError(rc, "hipMemGetInfo failed")
    return f.value, t.value


class ScfEngine:
    def __init__(self, prob: ScfProblem):
        L = load_library()
        self.L = L
        self.prob = prob
        self._H = np.asfortranarray(prob.H, dtype=complex)
        self._GL = np.asfortranarray(prob.gammaL, dtype=complex)
        self._GR = np.asfortranarray(prob.gammaR, dtype=complex)
        self._U = np.ascontiguousarray(prob.U_site, dtype=float)
        self._an = None if prob.atom_norb is None else np.ascontiguousarray(prob.atom_norb, dtype=np.int32)
        self.norb = self._H.shape[0]
        p = _Problem()
        p.norb = self.norb
        p.n_omega = prob.n_omega
        p.w_init = prob.w_init
        p.delta = prob.delta
        p.H = _ptr(self._H)
        p.gammaL = _ptr(self._GL)
        p.gammaR = _ptr(self._GR)
        p.U_site = _ptr(self._U)
        p.order = prob.order
        p.beta = prob.beta
        p.mu = prob.mu
        p.n_atoms = 0 if self._an is None else len(self._an)
        p.atom_norb = _ptr(self._an)
        p.alpha_mix = prob.alpha_mix
        p.nhist = prob.nhist
        p.maxscf = prob.maxscf
        p.tolf = prob.tolf
        p.tolf_order1 = prob.tolf_order1
        p.single_history = int(bool(prob.single_history))
        p.device = prob.device
        h = ctypes.c_void_p()
        rc = L.qs_create(ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise QsError(rc, "qs_create failed")
        self.h = h
        a = [ctypes.c_int32() for _ in range(4)]
        nv, nf = ctypes.c_int64(), ctypes.c_int32()
        L.qs_sizes(h, *[ctypes.byref(v) for v in a], ctypes.byref(nv), ctypes.byref(nf))
        self.nA, self.nS, self.nCL, self.nAs = [v.value for v in a]
        self.n_vec = nv.value
        self.Nf = nf.value

    def reset(self, prob: ScfProblem):
        """qs_reset_problem: the same handle (device memory, launch shape) for
        another system of the same sizes and active structure; QsError code 6 if
        it does not fit."""
        p, keep = _make_problem(prob)
        p.device = self.prob.device
        self._check(self.L.qs_reset_problem(self.h, ctypes.byref(p)))
        self.prob = prob
        self._H, self._GL, self._GR, self._U, self._an = keep

    def close(self):
        if getattr(self, "h", None):
            self.L.qs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise QsError(rc, self.L.qs_last_error(self.h).decode())

    def device_bytes(self):
        return self.L.qs_device_bytes(self.h)

    def solve_scf(self, V, fresh):
        r = _Result()
        m = max(self.prob.maxscf, 1)
        lf, lr = np.zeros(m), np.zeros(m)
        self._check(self.L.qs_solve_scf(self.h, float(V), int(bool(fresh)), ctypes.byref(r), _ptr(lf), _ptr(lr)))
        return ScfResult(bool(r.converged), r.n_iter, r.fmax, bool(r.runaway), bool(r.unphysical), r.i_worst,
                         r.n_worst, r.w_worst, lf[: r.n_iter].copy(), lr[: r.n_iter].copy())

    def current(self):
        u, d = ctypes.c_double(), ctypes.c_double()
        self._check(self.L.qs_current(self.h, ctypes.byref(u), ctypes.byref(d)))
        return u.value, d.value

    def g_less_obs(self):
        G = np.zeros((self.norb, self.norb), complex, order="F")
        self._check(self.L.qs_g_less_obs(self.h, _ptr(G)))
        return G

    def dosd(self):
        d = np.zeros((self.norb, self.prob.n_omega), order="F")
        self._check(self.L.qs_dosd(self.h, _ptr(d)))
        return d

    def trUD(self):
        u, d = np.zeros(self.prob.n_omega), np.zeros(self.prob.n_omega)
        self._check(self.L.qs_trUD(self.h, _ptr(u), _ptr(d)))
        return u, d

    def init2(self, V):
        x = np.zeros(self.n_vec)
        self._check(self.L.qs_init2(self.h, float(V), _ptr(x)))
        return x

    def residual2(self, V, x):
        x = np.ascontiguousarray(x, dtype=float)
        F = np.zeros(len(x))
        ab = ctypes.c_int32()
        self._check(self.L.qs_residual2(self.h, float(V), len(x), _ptr(x), _ptr(F), ctypes.byref(ab)))
        return F, bool(ab.value)

    def sigma(self):
        n = self.nA * self.nA * self.prob.n_omega
        r = np.zeros(n, complex)
        l = np.zeros(n, complex)
        s1 = np.zeros((self.norb, self.norb), complex, order="F")
        self._check(self.L.qs_get_sigma(self.h, _ptr(r), _ptr(l), _ptr(s1)))
        return r, l, s1

    def dyson_choice(self):
        v = ctypes.c_int32()
        ms = np.zeros(8)
        n = self.L.qs_dyson_choice(self.h, ctypes.byref(v), _ptr(ms), 8)
        return v.value, ms[:n].tolist()

    def set_precision(self, mode, switch_fmax=1e-5):
        """0 = FP64; 1 = mixed (complex-FP32 MFMA Dyson/Keldysh while max|F| >=
        switch_fmax, then FP64)"""
        self._check(self.L.qs_set_precision(self.h, int(mode), float(switch_fmax)))

    def set_step_mode(self, mode):
        """0 = host loop, 1 = device-controlled steps, 2 = device-controlled steps replayed from graphs"""
        self._check(self.L.qs_set_step_mode(self.h, int(mode)))

    def step_info(self):
        """(device-controlled steps apply, graph replays in the last solve)"""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.L.qs_step_info(self.h, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), b.value

    def precision_steps(self):
        """(FP32 residuals, FP64 residuals) of the last solve"""
        a, b = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.L.qs_precision_steps(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def dyson_shape_name(self):
        return self.L.qs_dyson_shape(self.h, None).decode()

    def dyson_uses_mfma(self):
        m = ctypes.c_int32()
        self.L.qs_dyson_shape(self.h, ctypes.byref(m))
        return bool(m.value)

    def profile(self, enable=True):
        self._check(self.L.qs_profile(self.h, int(bool(enable))))

    def profile_read(self):
        ms = np.zeros(32)
        cnt = np.zeros(32, dtype=np.int64)
        n = self.L.qs_profile_read(self.h, _ptr(ms), _ptr(cnt), 32)
        if n < 0:
            raise QsError(-n, "qs_profile_read")
        return {self.L.qs_stage_name(i).decode(): (ms[i], int(cnt[i])) for i in range(n)}

    def final_sweep(self, V):
        r = _Result()
        self._check(self.L.qs_final_sweep(self.h, float(V), ctypes.byref(r)))
        return r


_hip = None


class ScfGroup:
    """qs_group_*: several order-2 solves stepped together, with the
    Dyson sweeps of members that share a kernel instance in one launch.  Members
    are ScfEngine objects; add() starts a solve, step() runs one SCF step of every
    member and returns the solves that ended as (engine, ScfResult, status) with
    status 0 or the library's error code.  A finished engine's observables are
    read as after ScfEngine.solve_scf.  One thread drives a group."""

    def __init__(self, device=0):
        L = load_library()
        self.L = L
        g = ctypes.c_void_p()
        rc = L.qs_group_create(int(device), ctypes.byref(g))
        if rc != 0:
            raise QsError(rc, "qs_group_create failed")
        self.g = g
        self._members = {}

    def _check(self, rc):
        if rc != 0:
            raise QsError(rc, self.L.qs_group_last_error(self.g).decode())

    def add(self, eng: "ScfEngine", V, fresh=True):
        self._check(self.L.qs_group_add(self.g, eng.h, float(V), int(bool(fresh))))
        self._members[eng.h.value] = eng

    def __len__(self):
        return self.L.qs_group_members(self.g)

    def step(self):
        return self._finish(self.L.qs_group_step)

    def step_begin(self):
        """enqueue one step and return at once (step_end waits for it)"""
        self._check(self.L.qs_group_step_begin(self.g))

    def step_end(self):
        return self._finish(self.L.qs_group_step_end)

    def _finish(self, fn):
        nm = max(len(self), 1)
        done = (ctypes.c_void_p * nm)()
        res = (_Result * nm)()
        st = (ctypes.c_int32 * nm)()
        nd = ctypes.c_int32()
        self._check(fn(self.g, nm, done, res, st, ctypes.byref(nd)))
        out = []
        for k in range(nd.value):
            eng = self._members.pop(done[k])
            r = res[k]
            out.append((eng, ScfResult(bool(r.converged), r.n_iter, r.fmax, bool(r.runaway), bool(r.unphysical),
                                       r.i_worst, r.n_worst, r.w_worst), int(st[k])))
        return out

    def info(self):
        """(Dyson launches of the last step, members in grouped launches, ms of all
        steps' Dyson sweeps, steps)"""
        a, b, s = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        ms = ctypes.c_double()
        self._check(self.L.qs_group_info(self.g, ctypes.byref(a), ctypes.byref(b), ctypes.byref(ms), ctypes.byref(s)))
        return a.value, b.value, ms.value, s.value

    def close(self):
        if getattr(self, "g", None):
            self.L.qs_group_destroy(self.g)
            self.g = None
            self._members = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hip_runtime():
    """The HIP runtime libqs_hip.so itself runs on (found in this process's
    mappings once the library is loaded), for callers that hand over device
    memory.  A second runtime in the process (such as the one a framework
    bundles) cannot share the GPU with it."""
    global _hip
    if _hip is None:
        load_library()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    cand = line.split()[-1]
                    if path is None or "/opt/rocm" in cand:
                        path = cand
        if path is None:
            raise RuntimeError("libqs_hip.so is loaded but its HIP runtime is not mapped")
        H = ctypes.CDLL(path)
        P = ctypes.c_void_p
        H.hipMalloc.argtypes = [ctypes.POINTER(P), ctypes.c_size_t]
        H.hipFree.argtypes = [P]
        H.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
        H.hipSetDevice.argtypes = [ctypes.c_int]
        _hip = H
    return _hip


class DeviceBuffer:
    """Device memory from the library's own HIP runtime (hipMalloc / hipMemcpy)."""

    def __init__(self, nbytes, device=0):
        H = hip_runtime()
        self.H, self.nbytes, self.ptr = H, int(nbytes), ctypes.c_void_p()
        if H.hipSetDevice(int(device)) != 0 or H.hipMalloc(ctypes.byref(self.ptr), self.nbytes) != 0:
            raise QsError(4, "hipMalloc of %d bytes failed" % self.nbytes)

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        if self.H.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, 1) != 0:
            raise QsError(3, "hipMemcpy to device failed")

    def download(self, a):
        assert a.flags.c_contiguous and a.nbytes == self.nbytes
        if self.H.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, 2) != 0:
            raise QsError(3, "hipMemcpy to host failed")

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self.H.hipFree(self.ptr)
            self.ptr = None


class SanchoSelfEnergies:
    """SANCHO's interaction self-energies on one device.

    build(GR, GL, U_site, delta) takes G_in%r and G_in%L as (na, na, n_omega)
    complex numpy arrays and returns Sig2R, Sig2L, Sig2G and, at order 3,
    Sig3R, Sig3L, Sig3G, all in the reference's column-major layout.
    build_device() does the same on device memory (DeviceBuffer, or any
    address from the library's HIP runtime) with nothing copied to the host.
    """

    def __init__(self, na, n_omega, order, device=0):
        L = load_library()
        self.L = L
        self.na, self.n_omega, self.order, self.device = int(na), int(n_omega), int(order), int(device)
        h = ctypes.c_void_p()
        rc = L.qs_sancho_create(self.na, self.n_omega, self.order, self.device, ctypes.byref(h))
        if rc != 0:
            raise QsError(rc, "qs_sancho_create(na=%d, n_omega=%d, order=%d) failed" % (na, n_omega, order))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.qs_sancho_destroy(h)
            self.h = None

    @property
    def n_outputs(self):
        return 6 if self.order == 3 else 3

    def device_bytes(self):
        return int(self.L.qs_sancho_device_bytes(self.h))

    def profile(self, enable=True):
        rc = self.L.qs_sancho_profile(self.h, int(bool(enable)))
        if rc != 0:
            raise QsError(rc, self.L.qs_sancho_last_error(self.h).decode())

    def profile_read(self):
        """{"bubbles": (ms, builds), "self_energies": (ms, builds)} since profile(True)"""
        ms = np.zeros(2)
        cnt = np.zeros(2, dtype=np.int64)
        n = self.L.qs_sancho_profile_read(self.h, _ptr(ms), _ptr(cnt))
        if n < 0:
            raise QsError(-n, "qs_sancho_profile_read")
        return {"bubbles": (float(ms[0]), int(cnt[0])), "self_energies": (float(ms[1]), int(cnt[1]))}

    def _call(self, delta, U_site, gr, gl, outs, on_device):
        U = np.ascontiguousarray(U_site, dtype=np.float64)
        if U.shape != (self.na,):
            raise ValueError("U_site must have na = %d entries" % self.na)
        outs = list(outs) + [None] * (6 - len(outs))
        rc = self.L.qs_sancho_self_energies(self.h, float(delta), _ptr(U), gr, gl, *outs, on_device)
        if rc != 0:
            raise QsError(rc, self.L.qs_sancho_last_error(self.h).decode())

    def build(self, GR, GL, U_site, delta):
        shape = (self.na, self.na, self.n_omega)
        GR = np.asfortranarray(GR, dtype=complex)
        GL = np.asfortranarray(GL, dtype=complex)
        if GR.shape != shape or GL.shape != shape:
            raise ValueError("GR, GL must be %s" % (shape,))
        outs = [np.zeros(shape, complex, order="F") for _ in range(self.n_outputs)]
        self._call(delta, U_site, _ptr(GR), _ptr(GL), [_ptr(o) for o in outs], 0)
        return outs

    def build_device(self, gr, gl, outs, U_site, delta):
        """gr, gl, outs: DeviceBuffer objects (or device addresses) of na^2 n_omega complex128 each"""
        addr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else ctypes.c_void_p(int(b))
        if len(outs) != self.n_outputs:
            raise ValueError("order %d needs %d outputs" % (self.order, self.n_outputs))
        self._call(delta, U_site, addr(gr), addr(gl), [addr(o) for o in outs], 1)


def sancho_self_energies_static(na, n_omega, order, gr, gl, outs, U_site, delta, stream=None):
    """qs_sancho_self_energies_static: SANCHO's self-energies on device arrays with
    no handle and no device allocation (work arrays in the library's fixed buffer).
    gr, gl, outs: DeviceBuffer objects or device addresses of na^2 n_omega complex128."""
    L = load_library()
    addr = lambda b: b.ptr if isinstance(b, DeviceBuffer) else ctypes.c_void_p(int(b))
    n_out = 6 if order == 3 else 3
    if len(outs) != n_out:
        raise ValueError("order %d needs %d outputs" % (order, n_out))
    U = np.ascontiguousarray(U_site, dtype=np.float64)
    if U.shape != (na,):
        raise ValueError("U_site must have na = %d entries" % na)
    o = [addr(b) for b in outs] + [None] * (6 - n_out)
    rc = L.qs_sancho_self_energies_static(int(na), int(n_omega), int(order), float(delta), _ptr(U), addr(gr), addr(gl),
                                          *o, stream)
    if rc != 0:
        raise QsError(rc, "qs_sancho_self_energies_static(na=%d, n_omega=%d, order=%d)" % (na, n_omega, order))
