"""Host-side binding of liblesgo_hip.so (the HIP implementation) via ctypes.

Mirrors the reference's subroutine interface: `filt_da(f, dfdx, dfdy)`,
`ddx(dfdx, f)`, `press_stag_array()`, ... with the same argument meaning.  Arrays
may be numpy arrays (host memory, staged by the library) or torch tensors on the
GPU (device memory, used in place).  There is no CPU fallback: if the library is
missing or no GPU is visible, every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblesgo_hip.so")
_lib = None


class LesgoError(RuntimeError):
    """a non-zero lp_status; `code` holds it (None for errors raised on the host side)"""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def load():
    """Load liblesgo_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LesgoError(f"{LIB_PATH} is missing: build it with "
                         "`python -m population_balance_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7).  Loaded first, it also satisfies this library's dependency on
    # that soname; loaded after this library, torch would open a second runtime, which
    # then finds no GPU.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    vp, i32 = C.c_void_p, C.c_int32
    sig = {
        "lp_abi_version": ([], C.c_int),
        "lp_last_error": ([], C.c_char_p),
        "lp_grid_init": ([C.POINTER(abi.lp_grid), i32, i32, i32, C.c_double, C.c_double, C.c_double], C.c_int),
        "lp_config_default": ([C.POINTER(abi.lp_config)], None),
        "lp_create": ([C.POINTER(abi.lp_config), C.POINTER(vp)], C.c_int),
        "lp_destroy": ([vp], C.c_int),
        "lp_sync": ([vp], C.c_int),
        "lp_field_info": ([vp, i32, C.POINTER(vp), C.POINTER(C.c_size_t), C.POINTER(i32)], C.c_int),
        "lp_upload": ([vp, i32, vp, C.c_size_t], C.c_int),
        "lp_download": ([vp, i32, vp, C.c_size_t], C.c_int),
        "lp_set_time": ([vp, i32, i32], C.c_int),
        "lp_get_time": ([vp, C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "lp_init_wavenumber": ([vp, vp, vp, vp], C.c_int),
        "lp_test_filter_init": ([vp, C.c_double, vp], C.c_int),
        "lp_filt_da": ([vp, vp, vp, vp, i32], C.c_int),
        "lp_ddx": ([vp, vp, vp, i32], C.c_int),
        "lp_ddy": ([vp, vp, vp, i32], C.c_int),
        "lp_ddxy": ([vp, vp, vp, vp, i32], C.c_int),
        "lp_ddz_uv": ([vp, vp, vp], C.c_int),
        "lp_ddz_w": ([vp, vp, vp], C.c_int),
        "lp_padd": ([vp, vp, vp, i32], C.c_int),
        "lp_unpadd": ([vp, vp, vp, i32], C.c_int),
        "lp_dealias1": ([vp, vp, vp, i32], C.c_int),
        "lp_dealias2": ([vp, vp, vp, i32], C.c_int),
        "lp_test_filter": ([vp, vp, vp, i32], C.c_int),
        "lp_divstress_uv": ([vp, vp, vp, vp, vp], C.c_int),
        "lp_divstress_w": ([vp, vp, vp, vp, vp], C.c_int),
        "lp_tridag_array": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "lp_wallstress": ([vp], C.c_int),
        "lp_sgs_stag": ([vp], C.c_int),
        "lp_lagrange_sdep": ([vp], C.c_int),
        "lp_interpolag_sdep": ([vp], C.c_int),
        "lp_convec": ([vp], C.c_int),
        "lp_press_stag_array": ([vp], C.c_int),
        "lp_project": ([vp], C.c_int),
        "lp_timestep_conditions": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
        "lp_rmsdiv": ([vp, C.POINTER(C.c_double)], C.c_int),
        "lp_plane_stats": ([vp, C.POINTER(C.c_double)], C.c_int),
        "lp_spectrum": ([vp, i32, i32, C.POINTER(C.c_double)], C.c_int),
        "lp_energy": ([vp, C.POINTER(C.c_double)], C.c_int),
        "lp_obukhov": ([vp], C.c_int),
        "lp_obukhov_get": ([vp, C.POINTER(C.c_double)], C.c_int),
        "lp_theta_all_in_one": ([vp], C.c_int),
        "lp_tavg_reset": ([vp], C.c_int),
        "lp_tavg_compute": ([vp, C.c_double], C.c_int),
        "lp_tavg_get": ([vp, i32, vp, C.c_size_t, C.POINTER(C.c_double)], C.c_int),
        "lp_step": ([vp, i32], C.c_int),
        "lp_set_cfl_limit": ([vp, C.c_double], C.c_int),
        "lp_check_finite": ([vp], C.c_int),
        "lp_get_precision": ([vp, C.POINTER(i32)], C.c_int),
        "lp_set_frozen": ([vp, vp, vp], C.c_int),
        "lp_step_adjoint": ([vp, vp, vp, vp, vp], C.c_int),
        "lp_step_adjoint2": ([vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "lp_pollen_all_in_one2": ([vp], C.c_int),
        "lp_pollen_rhs_calc2": ([vp], C.c_int),
        "lp_calcbeta_pcon": ([vp], C.c_int),
        "lp_step_pollen2": ([vp], C.c_int),
        "lp_calculate_probability": ([vp, i32, C.c_double, C.c_double, vp], C.c_int),
        "lp_prof_enable": ([vp, i32], C.c_int),
        "lp_prof_read": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.lp_abi_version() != abi.LP_ABI_VERSION:
        raise LesgoError("ABI version mismatch between liblesgo_hip.so and abi.py")
    _lib = L
    return L


# the accumulators of lp_tavg_get, in order (then "pcon<b>", "pcon2<b>" per droplet bin)
TAVG_NAMES = ("u", "v", "w_uv", "u_w", "v_w", "w", "u2", "v2", "w2", "uv", "uw", "vw",
              "txx", "tyy", "tzz", "txy", "txz", "tyz", "p", "dudx", "dudy", "dudz_uv",
              "dvdx", "dvdy", "dvdz_uv", "dwdx_uv", "dwdy_uv", "dwdz", "tot_pc", "tot_pc2")


def _torch_ready(t):
    """the library's stream does not wait on torch's: work that torch has queued on a
    tensor (or on the memory a new tensor reuses) must be done before the library reads
    or writes it"""
    if t.is_cuda:
        import torch
        torch.cuda.current_stream(t.device).synchronize()


def _ptr(a):
    """device or host address of a numpy array / torch tensor."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise LesgoError("torch tensors must be contiguous")
        _torch_ready(a)
        return C.c_void_p(a.data_ptr())
    if not a.flags["C_CONTIGUOUS"]:
        raise LesgoError("numpy arrays must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def calculate_probability(diameter, sigma_s=15.5e-3, d_min=1e-7):
    """scalars_module_crossflow_mod.f90:4530-4597 on the host: beta[i, j] is the
    probability that a drop of bin j breaks into bin i."""
    L = load()
    d = np.ascontiguousarray(diameter, np.float64)
    beta = np.zeros((len(d), len(d)))
    rc = L.lp_calculate_probability(_ptr(d), len(d), float(sigma_s), float(d_min), _ptr(beta))
    if rc != abi.LP_OK:
        raise LesgoError(f"lp_calculate_probability failed ({rc}): {L.lp_last_error().decode()}")
    return beta


class Lesgo:
    """A context on the current HIP device (one stream, device-resident state)."""

    def __init__(self, cfg):
        self.L = load()
        self.cfg = cfg
        self.g = cfg.grid
        h = C.c_void_p()
        self._check(self.L.lp_create(C.byref(cfg), C.byref(h)), "lp_create")
        self.h = h
        self.fdtype = np.float32 if cfg.precision == abi.LP_MIXED else np.float64

    def _check(self, rc, what):
        if rc != abi.LP_OK:
            raise LesgoError(f"{what} failed ({rc}): {self.L.lp_last_error().decode()}", rc)

    def close(self):
        if getattr(self, "h", None):
            self.L.lp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    def field_info(self, name):
        p, b, e = C.c_void_p(), C.c_size_t(), C.c_int32()
        self._check(self.L.lp_field_info(self.h, abi.FIELD[name], C.byref(p), C.byref(b), C.byref(e)),
                    "lp_field_info")
        return p.value, b.value, e.value

    def field_dtype(self, name):
        return np.float64 if self.field_info(name)[2] == 8 else np.float32

    def upload(self, name, arr):
        _, nbytes, esz = self.field_info(name)
        a = np.ascontiguousarray(arr, dtype=np.float64 if esz == 8 else np.float32)
        if a.nbytes != nbytes:
            raise LesgoError(f"{name}: expected {nbytes} bytes, got {a.nbytes}")
        self._check(self.L.lp_upload(self.h, abi.FIELD[name], _ptr(a), a.nbytes), "lp_upload")

    def download(self, name):
        _, nbytes, esz = self.field_info(name)
        shape = abi.field_shape(self.g, name, self.cfg.drop.npcon)
        dt = np.float64 if esz == 8 else np.float32
        n = nbytes // esz
        out = np.empty(n, dt)
        self._check(self.L.lp_download(self.h, abi.FIELD[name], _ptr(out), nbytes), "lp_download")
        return out.reshape(shape) if int(np.prod(shape)) == n else out

    def get(self, name):
        """a field as a torch tensor on the device (a copy)"""
        import torch
        _, nbytes, esz = self.field_info(name)
        shape = abi.field_shape(self.g, name, self.cfg.drop.npcon)
        t = torch.empty(shape, dtype=torch.float64 if esz == 8 else torch.float32, device="cuda")
        if t.numel() * esz != nbytes:
            raise LesgoError(f"{name}: shape {shape} does not hold {nbytes} bytes")
        self._check(self.L.lp_download(self.h, abi.FIELD[name], _ptr(t), nbytes), "lp_download")
        return t

    def put(self, name, t):
        """set a field from a torch tensor on the device (converted to the field's type)"""
        import torch
        _, nbytes, esz = self.field_info(name)
        t = t.to(torch.float64 if esz == 8 else torch.float32).contiguous()
        if t.numel() * esz != nbytes:
            raise LesgoError(f"{name}: expected {nbytes} bytes, got {t.numel() * esz}")
        self._check(self.L.lp_upload(self.h, abi.FIELD[name], _ptr(t), nbytes), "lp_upload")

    def set_frozen(self, nut, ust):
        """frozen-coefficient steps with this Nu_t (volume) and wall u* (plane), device
        tensors of the field type; (None, None) restores the SGS model and the wall law"""
        self._frozen = (nut, ust)  # keep the tensors alive while the context uses them
        self._check(self.L.lp_set_frozen(self.h, _ptr(nut), _ptr(ust)), "set_frozen")

    def step_adjoint(self, nut, ust, ub, rb, cb=None, rcb=None):
        """the adjoint of one frozen-coefficient step at the context's state (binary64
        device tensors: nut, ust the step's coefficients; ub (3 volumes), rb (3 volumes)
        and, with droplets, cb, rcb (npcon volumes: PCon, RHS_PCon) updated in place from
        the adjoints of its outputs to those of its inputs)"""
        self._check(self.L.lp_step_adjoint2(self.h, _ptr(nut), _ptr(ust), _ptr(ub), _ptr(rb), _ptr(cb), _ptr(rcb)),
                    "step_adjoint")

    def set_time(self, jt, jt_total=None):
        self._check(self.L.lp_set_time(self.h, jt, jt if jt_total is None else jt_total), "lp_set_time")

    def time(self):
        a, b = C.c_int32(), C.c_int32()
        self._check(self.L.lp_get_time(self.h, C.byref(a), C.byref(b)), "lp_get_time")
        return a.value, b.value

    def sync(self):
        self._check(self.L.lp_sync(self.h), "lp_sync")

    # ---- spectral primitives (reference argument order) ----
    def init_wavenumber(self):
        g = self.g
        kx, ky, k2 = (np.zeros((g.ny, g.lh)) for _ in range(3))
        self._check(self.L.lp_init_wavenumber(self.h, _ptr(kx), _ptr(ky), _ptr(k2)), "init_wavenumber")
        return kx, ky, k2

    def test_filter_init(self, alpha):
        G = np.zeros((self.g.ny, self.g.lh))
        self._check(self.L.lp_test_filter_init(self.h, float(alpha), _ptr(G)), "test_filter_init")
        return G

    def _nlev(self, a):
        return int(a.shape[0]) if len(a.shape) == 3 else 1

    def filt_da(self, f, dfdx, dfdy):
        """in place on f (filtered), writes dfdx, dfdy."""
        self._check(self.L.lp_filt_da(self.h, _ptr(f), _ptr(dfdx), _ptr(dfdy), self._nlev(f)), "filt_da")

    def ddx(self, dfdx, f):
        self._check(self.L.lp_ddx(self.h, _ptr(dfdx), _ptr(f), self._nlev(f)), "ddx")

    def ddy(self, dfdy, f):
        self._check(self.L.lp_ddy(self.h, _ptr(dfdy), _ptr(f), self._nlev(f)), "ddy")

    def ddxy(self, dfdx, dfdy, f):
        self._check(self.L.lp_ddxy(self.h, _ptr(dfdx), _ptr(dfdy), _ptr(f), self._nlev(f)), "ddxy")

    def ddz_uv(self, dfdz, f):
        self._check(self.L.lp_ddz_uv(self.h, _ptr(dfdz), _ptr(f)), "ddz_uv")

    def ddz_w(self, dfdz, f):
        self._check(self.L.lp_ddz_w(self.h, _ptr(dfdz), _ptr(f)), "ddz_w")

    def padd(self, u_big, u):
        self._check(self.L.lp_padd(self.h, _ptr(u_big), _ptr(u), int(u.shape[0])), "padd")

    def unpadd(self, cc, cc_big):
        self._check(self.L.lp_unpadd(self.h, _ptr(cc), _ptr(cc_big), int(cc.shape[0])), "unpadd")

    def dealias1(self, u, u_big):
        self._check(self.L.lp_dealias1(self.h, _ptr(u), _ptr(u_big), self._nlev(u)), "dealias1")

    def dealias2(self, u, u_big):
        self._check(self.L.lp_dealias2(self.h, _ptr(u), _ptr(u_big), self._nlev(u)), "dealias2")

    def test_filter(self, f, G):
        self._check(self.L.lp_test_filter(self.h, _ptr(f), _ptr(G), self._nlev(f)), "test_filter")

    def divstress_uv(self, divt, tx, ty, tz):
        self._check(self.L.lp_divstress_uv(self.h, _ptr(divt), _ptr(tx), _ptr(ty), _ptr(tz)), "divstress_uv")

    def divstress_w(self, divt, tx, ty, tz):
        self._check(self.L.lp_divstress_w(self.h, _ptr(divt), _ptr(tx), _ptr(ty), _ptr(tz)), "divstress_w")

    def tridag_array(self, a, b, c, r, u):
        self._check(self.L.lp_tridag_array(self.h, _ptr(a), _ptr(b), _ptr(c), _ptr(r), _ptr(u)), "tridag_array")

    # ---- routines on the state ----
    def wallstress(self):
        self._check(self.L.lp_wallstress(self.h), "wallstress")

    def sgs_stag(self):
        self._check(self.L.lp_sgs_stag(self.h), "sgs_stag")

    def lagrange_sdep(self):
        self._check(self.L.lp_lagrange_sdep(self.h), "lagrange_sdep")

    def interpolag_sdep(self):
        self._check(self.L.lp_interpolag_sdep(self.h), "interpolag_sdep")

    def convec(self):
        self._check(self.L.lp_convec(self.h), "convec")

    def press_stag_array(self):
        self._check(self.L.lp_press_stag_array(self.h), "press_stag_array")

    def project(self):
        self._check(self.L.lp_project(self.h), "project")

    def timestep_conditions(self):
        a, b = C.c_double(), C.c_double()
        self._check(self.L.lp_timestep_conditions(self.h, C.byref(a), C.byref(b)), "timestep_conditions")
        return a.value, b.value

    def set_cfl_limit(self, cfl_max):
        """lp_step raises (LP_ERR_CFL) when the CFL after its steps exceeds cfl_max (0: off)."""
        self._check(self.L.lp_set_cfl_limit(self.h, float(cfl_max)), "set_cfl_limit")

    def check_finite(self):
        """raises LesgoError (LP_ERR_NAN) naming the first state field with a non-finite value."""
        self._check(self.L.lp_check_finite(self.h), "check_finite")

    def precision(self):
        a = C.c_int32()
        self._check(self.L.lp_get_precision(self.h, C.byref(a)), "get_precision")
        return a.value

    def rmsdiv(self):
        a = C.c_double()
        self._check(self.L.lp_rmsdiv(self.h, C.byref(a)), "rmsdiv")
        return a.value

    def plane_stats(self):
        """(nz-1, 10) per-level statistics of avg_stats (see lp_plane_stats)."""
        out = np.zeros((self.g.nz - 1, 10))
        self._check(self.L.lp_plane_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "plane_stats")
        return out

    def spectrum(self, name, bin=0):
        """(nz-1, nx/2) streamwise energy spectrum of every level (see lp_spectrum); avg_stats'
        E11(k1, z)/z is row k divided by (k + 1/2) dz."""
        out = np.zeros((self.g.nz - 1, self.g.nx // 2))
        rc = self.L.lp_spectrum(self.h, abi.FIELD[name], bin, out.ctypes.data_as(C.POINTER(C.c_double)))
        self._check(rc, "spectrum")
        return out

    def energy(self):
        """kea of energy.f90: the mean of (u^2 + v^2 + w_uv^2)/2 over levels 1..nz-1."""
        a = C.c_double()
        self._check(self.L.lp_energy(self.h, C.byref(a)), "energy")
        return a.value

    def obukhov(self):
        """obukhov (see lp_obukhov): the stability functions, psi_m and the u* plane."""
        self._check(self.L.lp_obukhov(self.h), "obukhov")

    def obukhov_get(self):
        """the last obukhov call: psi_m, psi_h, phi_m, phi_h, u*, L, <sgs_t3(1)>, <theta(1)>"""
        out = np.zeros(8)
        self._check(self.L.lp_obukhov_get(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "obukhov_get")
        return out

    def theta_all_in_one(self):
        """theta_all_in_one (see lp_theta_all_in_one): one temperature step."""
        self._check(self.L.lp_theta_all_in_one(self.h), "theta_all_in_one")

    def tavg_reset(self):
        """tavg_init: zero the running time averages and their total time."""
        self._check(self.L.lp_tavg_reset(self.h), "tavg_reset")

    def tavg_compute(self, tavg_dt):
        """tavg_compute: add tavg_dt times each averaged quantity of the current state."""
        self._check(self.L.lp_tavg_compute(self.h, float(tavg_dt)), "tavg_compute")

    def tavg_get(self, q):
        """(accumulator q as a binary64 (nz, ny, ld) array, tavg_total_time); the
        quantities' order is that of lp_tavg_get (TAVG_NAMES)."""
        q = TAVG_NAMES.index(q) if isinstance(q, str) else int(q)
        out = np.zeros((self.g.nz, self.g.ny, self.g.ld))
        t = C.c_double()
        rc = self.L.lp_tavg_get(self.h, q, out.ctypes.data, out.nbytes, C.byref(t))
        self._check(rc, "tavg_get")
        return out, t.value

    def pollen_all_in_one2(self):
        self._check(self.L.lp_pollen_all_in_one2(self.h), "pollen_all_in_one2")

    def pollen_rhs_calc2(self):
        self._check(self.L.lp_pollen_rhs_calc2(self.h), "pollen_rhs_calc2")

    def calcbeta_pcon(self):
        self._check(self.L.lp_calcbeta_pcon(self.h), "calcbeta_pcon")

    def step_pollen2(self):
        self._check(self.L.lp_step_pollen2(self.h), "step_pollen2")

    def step(self, n=1):
        self._check(self.L.lp_step(self.h, int(n)), "step")

    # ---- measurement ----
    def prof_enable(self, kernel_class):
        self._check(self.L.lp_prof_enable(self.h, int(kernel_class)), "prof_enable")

    def prof_read(self):
        ms, n, b = C.c_double(), C.c_int64(), C.c_double()
        self._check(self.L.lp_prof_read(self.h, C.byref(ms), C.byref(n), C.byref(b)), "prof_read")
        return ms.value, n.value, b.value
