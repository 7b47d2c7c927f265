"""yahmc_amd — MI355X-native fermion hot path of YaHMC.

Python host mirror of the reference's call surface (include/dirac.h,
include/inverter.h, include/linalg.h, include/clover.h) over the C ABI in
include/yahmc.h.  Every compute call runs in libyahmc.so's HIP kernels; if
the library is missing, import of the native layer raises immediately —
there is no CPU fallback.

Host arrays use the reference's layouts: spinor fields (V, 8*DR) float64
(Spinor re[DR][4] then im[DR][4], EO site order), fundamental links
(4V, 2*NC*NC) float64 (suNg re-plane then im-plane, index 4*id+mu).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libyahmc.so")

YH_OK = 0
TUNE_XCD_CHUNK, TUNE_NET_PREFETCH, TUNE_GRAPHS, TUNE_HALF_LINKS, TUNE_RECON12_F64, TUNE_MULTI_RHS = 1, 2, 3, 4, 5, 6
TUNE_LEARNED_GCR = 7
TUNE_HOP_STAGE = 8
TUNE_FORCE_ORDER = 9
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "bad state", -4: "not converged", -5: "non-finite",
          -6: "unsupported"}

REPR = {"fundamental": 0, "adjoint": 1, "symmetric": 2, "antisymmetric": 3}
BC = {"periodic": 0, "antiperiodic": 1, "open": 2, "sf": 3}
PREC = {"f64": 0, "f32": 1, "f16": 2}
MVM = {"Dphi": 0, "Hphi": 1, "Hphi_sq": 2}
EVEN, ODD, BOTH = 1, 2, 3
RECON_18, RECON_12 = 0, 1


class YahmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class ConfigInfo(ctypes.Structure):
    """yh_config_info: the fields of header_t (include/gaugefield.h:24-36)."""
    _fields_ = [("ident", ctypes.c_int), ("version", ctypes.c_int), ("dims", ctypes.c_int * 4), ("nc", ctypes.c_int),
                ("repr", ctypes.c_int), ("num_cnfg", ctypes.c_int), ("num_accept", ctypes.c_int),
                ("plaquette", ctypes.c_double)]


CFG = {"native": 0, "hirep": 1}


class Rational(ctypes.Structure):
    """yh_rational: the fields of rational_app (include/rational.h:7-18)."""
    _fields_ = [("n", ctypes.c_int), ("d", ctypes.c_int), ("order", ctypes.c_int), ("prec", ctypes.c_double),
                ("min", ctypes.c_double), ("max", ctypes.c_double), ("a0", ctypes.c_double),
                ("pole", ctypes.c_double * 48), ("root", ctypes.c_double * 48), ("residue", ctypes.c_double * 48)]

    def __call__(self, x):
        return lib().yh_rational_eval(ctypes.byref(self), x)


class _Lattice(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 4), ("nc", ctypes.c_int), ("repr", ctypes.c_int), ("bc", ctypes.c_int),
                ("csw", ctypes.c_double), ("cf", ctypes.c_double), ("device", ctypes.c_int)]


class InvPar(ctypes.Structure):
    """yh_inv_par: mirrors inv_par (include/inverter.h:8-16)."""
    _fields_ = [("mass", ctypes.c_double), ("prec", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("mvm", ctypes.c_int), ("sloppy_prec", ctypes.c_int), ("reliable_every", ctypes.c_int),
                ("delta", ctypes.c_double), ("log", ctypes.c_int)]


_lib = None

_VP = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)

_SIGS = {
    "yh_create": ([ctypes.POINTER(_Lattice), ctypes.POINTER(_VP)], ctypes.c_int),
    "yh_destroy": ([_VP], ctypes.c_int),
    "yh_last_error": ([], ctypes.c_char_p),
    "yh_version": ([], ctypes.c_int),
    "yh_repr_dim": ([_VP], ctypes.c_int),
    "yh_volume": ([_VP], ctypes.c_long),
    "yh_set_gauge": ([_VP, _D], ctypes.c_int),
    "yh_get_gauge_repr": ([_VP, _D], ctypes.c_int),
    "yh_get_clover_term": ([_VP, _D], ctypes.c_int),
    "yh_clover_logdet": ([_VP, ctypes.c_double, _D], ctypes.c_int),
    "yh_field_create": ([_VP, ctypes.c_int, ctypes.POINTER(_VP)], ctypes.c_int),
    "yh_field_destroy": ([_VP], ctypes.c_int),
    "yh_field_upload": ([_VP, _VP, _D, ctypes.c_int], ctypes.c_int),
    "yh_field_download": ([_VP, _VP, _D, ctypes.c_int], ctypes.c_int),
    "yh_field_zero": ([_VP, _VP, ctypes.c_int], ctypes.c_int),
    "yh_field_copy": ([_VP, _VP, _VP, ctypes.c_int], ctypes.c_int),
    "yh_mvm_reset": ([_VP], None),
    "yh_mvm_get": ([_VP], ctypes.c_int),
    "yh_Dphi_oe": ([_VP, _VP, _VP], ctypes.c_int),
    "yh_Dphi_eo": ([_VP, _VP, _VP], ctypes.c_int),
    "yh_Dphi_ee": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_Dphi_oo": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_Dphi_oo_inv": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_Dphi": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_Hphi": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_Hphi_sq": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_spinor_sqnorm": ([_VP, _VP, _D], ctypes.c_int),
    "yh_spinor_product": ([_VP, _VP, _VP, _D], ctypes.c_int),
    "yh_spinor_product_re": ([_VP, _VP, _VP, _D], ctypes.c_int),
    "yh_spinor_g5": ([_VP, _VP], ctypes.c_int),
    "yh_spinor_mulr_add_assign": ([_VP, _VP, ctypes.c_double, _VP], ctypes.c_int),
    "yh_spinor_mulc_add_assign": ([_VP, _VP, ctypes.c_double, ctypes.c_double, _VP], ctypes.c_int),
    "yh_cg": ([_VP, ctypes.POINTER(InvPar), _VP, _VP, _PI], ctypes.c_int),
    "yh_cg_noalloc": ([_VP, ctypes.POINTER(InvPar), _VP, _VP, _PI], ctypes.c_int),
    "yh_bicgstab_mixed": ([_VP, ctypes.POINTER(InvPar), _VP, _VP, _PI, _PI], ctypes.c_int),
    "yh_gcr_net": ([_VP, ctypes.POINTER(InvPar), _VP, ctypes.c_int, _VP, _VP, _PI, _PI], ctypes.c_int),
    "yh_bicgstab": ([_VP, ctypes.POINTER(InvPar), _VP, _VP, _PI], ctypes.c_int),
    "yh_cg_mshift": ([_VP, ctypes.POINTER(InvPar), ctypes.c_int, _D, ctypes.POINTER(_VP), _VP, _PI], ctypes.c_int),
    "yh_cg_mixed": ([_VP, ctypes.POINTER(InvPar), _VP, _VP, _PI, _PI], ctypes.c_int),
    "yh_cg_mshift_batched": ([_VP, ctypes.POINTER(InvPar), ctypes.c_int, ctypes.c_int, _D, ctypes.POINTER(_VP),
                              ctypes.POINTER(_VP), _PI], ctypes.c_int),
    "yh_rational_zolotarev": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.POINTER(Rational)], ctypes.c_int),
    "yh_rational_remez": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                           ctypes.POINTER(Rational)], ctypes.c_int),
    "yh_rational_find": ([ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                          ctypes.POINTER(Rational)], ctypes.c_int),
    "yh_rational_eval": ([ctypes.POINTER(Rational), ctypes.c_double], ctypes.c_double),
    "yh_rational_calc": ([_VP, ctypes.POINTER(Rational), ctypes.POINTER(InvPar), _VP, _VP, _PI], ctypes.c_int),
    "yh_jacobi_sn": ([ctypes.c_double, ctypes.c_double], ctypes.c_double),
    "yh_ellip_k": ([ctypes.c_double], ctypes.c_double),
    "yh_set_momenta": ([_VP, _D], ctypes.c_int),
    "yh_get_momenta": ([_VP, _D], ctypes.c_int),
    "yh_get_clover_force": ([_VP, _D], ctypes.c_int),
    "yh_get_gauge": ([_VP, _D], ctypes.c_int),
    "yh_represent_gauge_field": ([_VP], ctypes.c_int),
    "yh_wilson_gauge_force": ([_VP, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "yh_improved_gauge_force": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double],
                                ctypes.c_int),
    "yh_avr_wilson": ([_VP, ctypes.c_int, ctypes.c_int, _D], ctypes.c_int),
    "yh_update_links": ([_VP, ctypes.c_double], ctypes.c_int),
    "yh_reunitarize": ([_VP], ctypes.c_int),
    "yh_apply_bcs_on_gauge_field": ([_VP], ctypes.c_int),
    "yh_plaquette": ([_VP, _D], ctypes.c_int),
    "yh_momenta_energy": ([_VP, _D], ctypes.c_int),
    "yh_begin_fermion_force": ([_VP], ctypes.c_int),
    "yh_end_fermion_force": ([_VP, ctypes.c_double], ctypes.c_int),
    "yh_clover_fermion_force": ([_VP, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_wilson_fermion_force": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "yh_wilson_fermion_force_q": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, _VP], ctypes.c_int),
    "yh_clover_logdet_force": ([_VP, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "yh_bench_hop": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP, _VP, _D], ctypes.c_int),
    "yh_bench_copy": ([_VP, ctypes.c_long, ctypes.c_int, _D], ctypes.c_int),
    "yh_synchronize": ([_VP], ctypes.c_int),
    "yh_net_create": ([_VP, ctypes.c_int, _PI, ctypes.POINTER(_VP)], ctypes.c_int),
    "yh_net_destroy": ([_VP], ctypes.c_int),
    "yh_net_num_weights": ([_VP], ctypes.c_long),
    "yh_net_set_weights": ([_VP, _D], ctypes.c_int),
    "yh_net_get_weights": ([_VP, _D], ctypes.c_int),
    "yh_net_apply": ([_VP, ctypes.c_int, _VP, _VP], ctypes.c_int),
    "yh_net_precondition": ([_VP, _VP, _VP], ctypes.c_int),
    "yh_net_loss_grad": ([_VP, ctypes.c_int, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _D, _D], ctypes.c_int),
    "yh_net_adam_step": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int],
                         ctypes.c_int),
    "yh_net_layer_host": ([_VP, ctypes.c_int, ctypes.c_int, _D, _D], ctypes.c_int),
    "yh_net_wgrad_host": ([_VP, ctypes.c_int, _D, _D, _D], ctypes.c_int),
    "yh_cg_net": ([_VP, ctypes.c_void_p, _VP, _VP, _VP, _PI, _PI], ctypes.c_int),
    "yh_cg_learned": ([_VP, ctypes.c_void_p, _VP, _VP, _VP, _VP, _PI, _PI, _PI], ctypes.c_int),
    "yh_tune": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "yh_autotune": ([_VP, ctypes.c_int, _PI], ctypes.c_int),
    "yh_config_write": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, _D], ctypes.c_int),
    "yh_config_read": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, _D], ctypes.c_int),
    "yh_calc_propagator": ([_VP, ctypes.c_void_p, _VP, _VP, _PI], ctypes.c_int),
    "yh_meson_contract": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_VP), ctypes.c_int, _D], ctypes.c_int),
    "yh_minres": ([_VP, ctypes.c_void_p, _VP, _VP, _PI], ctypes.c_int),
    "yh_auto_inverter": ([_VP, ctypes.c_void_p, _VP, _VP, _PI], ctypes.c_int),
    "yh_mre_create": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_VP)], ctypes.c_int),
    "yh_mre_destroy": ([_VP], ctypes.c_int),
    "yh_mre_reset": ([_VP], ctypes.c_int),
    "yh_mre_store": ([_VP, _VP], ctypes.c_int),
    "yh_mre_guess": ([_VP, _VP, _VP], ctypes.c_int),
}


def exported_symbols():
    return sorted(_SIGS)


def config_write(path, links, dims, nc, repr=0, plaquette=0.0, num_cnfg=0, num_accept=0, format="native"):
    """gauge_field_save / gauge_field_save_hirep: links (4V, 2 NC^2) in the
    gfield_fund layout (EO order)."""
    info = ConfigInfo()
    for i in range(4):
        info.dims[i] = int(dims[i])
    info.nc, info.repr, info.num_cnfg, info.num_accept, info.plaquette = nc, repr, num_cnfg, num_accept, plaquette
    links = np.ascontiguousarray(links, dtype=np.float64)
    assert links.size == 4 * int(np.prod(dims)) * 2 * nc * nc
    _check(lib().yh_config_write(str(path).encode(), CFG[format], ctypes.byref(info), _dp(links)))


def config_read(path, format="native"):
    """gauge_field_load / gauge_field_load_hirep: returns (info dict, links)."""
    info = ConfigInfo()
    _check(lib().yh_config_read(str(path).encode(), CFG[format], ctypes.byref(info), None))
    dims = tuple(info.dims)
    nc = info.nc
    links = np.zeros((4 * int(np.prod(dims)), 2 * nc * nc))
    _check(lib().yh_config_read(str(path).encode(), CFG[format], ctypes.byref(info), _dp(links)))
    return {"dims": dims, "nc": nc, "repr": info.repr, "num_cnfg": info.num_cnfg, "num_accept": info.num_accept,
            "plaquette": info.plaquette}, links


def rational_zolotarev(order, n, d, xmin, xmax):
    """Zolotarev approximation of x^(n/d), n/d = +-1/2 (host computation)."""
    r = Rational()
    _check(lib().yh_rational_zolotarev(order, n, d, xmin, xmax, ctypes.byref(r)))
    return r


def rational_remez(order, n, d, xmin, xmax):
    """Minimax relative-error approximation of x^(n/d), 0 < |n/d| < 1, by
    the Remez exchange (host computation)."""
    r = Rational()
    _check(lib().yh_rational_remez(order, n, d, xmin, xmax, ctypes.byref(r)))
    return r


def rational_find(n, d, prec, xmin, xmax):
    """Lowest-order approximation of x^(n/d) with max relative error < prec
    on [xmin, xmax] (mirrors rational_find, library/rational.cpp:56-132)."""
    r = Rational()
    _check(lib().yh_rational_find(n, d, prec, xmin, xmax, ctypes.byref(r)))
    return r


def lib():
    """Load libyahmc.so (built in-tree by yahmc_amd/build.py).  Raises if it
    is missing: the HIP path is the only implementation."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: build it with `python -m yahmc_amd.build` "
                              "(there is no CPU fallback)")
        L = ctypes.CDLL(LIBPATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != YH_OK:
        raise YahmcError(rc, lib().yh_last_error().decode())


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class Field:
    """A device spinor field (both parities) in HBM."""

    def __init__(self, lat, prec="f64"):
        self.lat = lat
        self.prec = PREC[prec] if isinstance(prec, str) else prec
        h = _VP()
        _check(lib().yh_field_create(lat.ctx, self.prec, ctypes.byref(h)))
        self.h = h

    def upload(self, host, mask=BOTH):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.shape == (self.lat.V, 8 * self.lat.dr)
        _check(lib().yh_field_upload(self.lat.ctx, self.h, _dp(host), mask))
        return self

    def download(self, mask=BOTH):
        out = np.zeros((self.lat.V, 8 * self.lat.dr))
        _check(lib().yh_field_download(self.lat.ctx, self.h, _dp(out), mask))
        return out

    def zero(self, mask=BOTH):
        _check(lib().yh_field_zero(self.lat.ctx, self.h, mask))

    def copy_from(self, other, mask=BOTH):
        _check(lib().yh_field_copy(self.lat.ctx, self.h, other.h, mask))

    def free(self):
        if self.h:
            lib().yh_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Lattice:
    """One lattice context on one GPU (geometry_init + repr_init +
    clover_init + bcs_init of the reference)."""

    def __init__(self, dims, nc=3, repr="fundamental", bc="antiperiodic", csw=1.0, cf=1.0, device=0):
        lt = _Lattice()
        for i in range(4):
            lt.dims[i] = int(dims[i])
        lt.nc = nc
        lt.repr = REPR[repr] if isinstance(repr, str) else repr
        lt.bc = BC[bc] if isinstance(bc, str) else bc
        lt.csw = csw
        lt.cf = cf
        lt.device = device
        h = _VP()
        _check(lib().yh_create(ctypes.byref(lt), ctypes.byref(h)))
        self.ctx = h
        self.dims = tuple(int(d) for d in dims)
        self.nc = nc
        self.csw = float(csw)
        self.bc = lt.bc
        self.repr_id = lt.repr
        self.dr = lib().yh_repr_dim(h)
        self.V = int(np.prod(self.dims))
        self.VH = self.V // 2

    def close(self):
        if self.ctx:
            lib().yh_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- gauge ---------------------------------------------------------
    def set_gauge(self, links):
        links = np.ascontiguousarray(links, dtype=np.float64)
        assert links.size == 4 * self.V * 2 * self.nc ** 2
        _check(lib().yh_set_gauge(self.ctx, _dp(links)))

    def gauge_repr(self):
        out = np.zeros((4 * self.V, 2 * self.dr ** 2))
        _check(lib().yh_get_gauge_repr(self.ctx, _dp(out)))
        return out

    def clover_term(self):
        out = np.zeros((4 * self.V, 2 * self.dr ** 2))
        _check(lib().yh_get_clover_term(self.ctx, _dp(out)))
        return out

    def clover_logdet(self, mass):
        out = np.zeros(1)
        _check(lib().yh_clover_logdet(self.ctx, mass, _dp(out)))
        return out[0]

    def field(self, prec="f64"):
        return Field(self, prec)

    # ---- operators (include/dirac.h) ----------------------------------
    def mvm_reset(self):
        lib().yh_mvm_reset(self.ctx)

    def mvm_get(self):
        return lib().yh_mvm_get(self.ctx)

    def Dphi_oe(self, d, s):
        _check(lib().yh_Dphi_oe(self.ctx, d.h, s.h))

    def Dphi_eo(self, d, s):
        _check(lib().yh_Dphi_eo(self.ctx, d.h, s.h))

    def Dphi_ee(self, m, d, s):
        _check(lib().yh_Dphi_ee(self.ctx, m, d.h, s.h))

    def Dphi_oo(self, m, d, s):
        _check(lib().yh_Dphi_oo(self.ctx, m, d.h, s.h))

    def Dphi_oo_inv(self, m, d, s):
        _check(lib().yh_Dphi_oo_inv(self.ctx, m, d.h, s.h))

    def Dphi(self, m, d, s):
        _check(lib().yh_Dphi(self.ctx, m, d.h, s.h))

    def Hphi(self, m, d, s):
        _check(lib().yh_Hphi(self.ctx, m, d.h, s.h))

    def Hphi_sq(self, m, d, s):
        _check(lib().yh_Hphi_sq(self.ctx, m, d.h, s.h))

    # ---- linalg (include/linalg.h) -----------------------------------
    def spinor_sqnorm(self, a):
        out = np.zeros(1)
        _check(lib().yh_spinor_sqnorm(self.ctx, a.h, _dp(out)))
        return out[0]

    def spinor_product(self, a, b):
        out = np.zeros(2)
        _check(lib().yh_spinor_product(self.ctx, a.h, b.h, _dp(out)))
        return complex(out[0], out[1])

    def spinor_product_re(self, a, b):
        out = np.zeros(1)
        _check(lib().yh_spinor_product_re(self.ctx, a.h, b.h, _dp(out)))
        return out[0]

    def spinor_g5(self, a):
        _check(lib().yh_spinor_g5(self.ctx, a.h))

    def spinor_mulr_add_assign(self, a, r, b):
        _check(lib().yh_spinor_mulr_add_assign(self.ctx, a.h, r, b.h))

    def spinor_mulc_add_assign(self, a, c, b):
        _check(lib().yh_spinor_mulc_add_assign(self.ctx, a.h, c.real, c.imag, b.h))

    # ---- solvers (include/inverter.h) --------------------------------
    @staticmethod
    def inv_par(mass, prec, mvm="Hphi_sq", max_iter=0, sloppy="f32", reliable_every=0, delta=0.1, log=0):
        p = InvPar()
        p.mass = mass
        p.prec = prec
        p.max_iter = max_iter
        p.mvm = MVM[mvm] if isinstance(mvm, str) else mvm
        p.sloppy_prec = PREC[sloppy] if isinstance(sloppy, str) else sloppy
        p.reliable_every = reliable_every
        p.delta = delta
        p.log = log
        return p

    def cg_inverter(self, par, d, s):
        it = ctypes.c_int(0)
        _check(lib().yh_cg(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def cg_noalloc(self, par, d, s):
        """yh_cg without allocation (temporaries in the library's per-GPU
        scratch array; see include/yahmc.h)."""
        it = ctypes.c_int(0)
        _check(lib().yh_cg_noalloc(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def bicgstab_inverter(self, par, d, s):
        it = ctypes.c_int(0)
        _check(lib().yh_bicgstab(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def cg_mshift(self, par, shifts, dsts, s):
        shifts = np.ascontiguousarray(shifts, dtype=np.float64)
        arr = (_VP * len(dsts))(*[d.h for d in dsts])
        it = ctypes.c_int(0)
        _check(lib().yh_cg_mshift(self.ctx, ctypes.byref(par), len(dsts), _dp(shifts), arr, s.h, ctypes.byref(it)))
        return it.value

    def cg_mshift_batched(self, par, shifts, dsts, srcs):
        """dsts: list (per rhs) of lists (per shift) of fields."""
        shifts = np.ascontiguousarray(shifts, dtype=np.float64)
        flat = [f.h for row in dsts for f in row]
        arr = (_VP * len(flat))(*flat)
        sarr = (_VP * len(srcs))(*[s.h for s in srcs])
        its = (ctypes.c_int * len(srcs))()
        _check(lib().yh_cg_mshift_batched(self.ctx, ctypes.byref(par), len(srcs), len(shifts), _dp(shifts), arr, sarr,
                                          its))
        return list(its)

    def rational_calc(self, rat, par, d, s):
        it = ctypes.c_int(0)
        _check(lib().yh_rational_calc(self.ctx, ctypes.byref(rat), ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def calc_propagator(self, par, d, s):
        """calc_propagator (observables/propagator.cpp:75-166): full-lattice
        D d = s; returns the CG iterations of the even solve."""
        it = ctypes.c_int(0)
        _check(lib().yh_calc_propagator(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def meson_contract(self, method, props, tau):
        """16 x T correlators of one source slice (mesons.cpp:55-147)."""
        arr = (_VP * len(props))(*[p.h for p in props])
        out = np.zeros((16, self.dims[0]))
        _check(lib().yh_meson_contract(self.ctx, method, len(props), arr, tau, _dp(out)))
        return out

    def save_config(self, path, format="native", num_cnfg=0, num_accept=0):
        """gauge_field_save (gaugefield.cpp:44-115) of the device links."""
        config_write(path, self.get_gauge(), self.dims, self.nc, self.repr_id, self.plaquette(), num_cnfg, num_accept,
                     format)

    def load_config(self, path, format="native"):
        """gauge_field_load (gaugefield.cpp:117-179): checks the header
        against this lattice, uploads and represents the links; returns the
        header and |stored plaquette - recomputed plaquette|."""
        info, links = config_read(path, format)
        bad = tuple(info["dims"]) != self.dims or info["nc"] != self.nc or \
            (format == "native" and info["repr"] != self.repr_id)
        if bad:
            raise YahmcError(-1, f"configuration {path} is not consistent with this lattice")
        self.set_gauge(links)
        return info, abs(info["plaquette"] - self.plaquette())

    def bicgstab_mixed(self, par, d, s):
        """BiCGstab on FP32 fields inside FP64 reliable updates; returns
        (2 x iterations, reliable updates)."""
        it = ctypes.c_int(0)
        up = ctypes.c_int(0)
        _check(lib().yh_bicgstab_mixed(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it), ctypes.byref(up)))
        return it.value, up.value

    def minres_inverter(self, par, d, s):
        """minres_inverter (library/minres.cpp:12-118); d holds the guess."""
        it = ctypes.c_int(0)
        _check(lib().yh_minres(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def auto_inverter(self, par, d, s):
        """auto_inverter (library/inverter.cpp:9-55)."""
        it = ctypes.c_int(0)
        _check(lib().yh_auto_inverter(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it)))
        return it.value

    def mre(self, max_past, mvm="Dphi", mass=0.0):
        return MRE(self, max_past, mvm, mass)

    def cg_mixed(self, par, d, s):
        it = ctypes.c_int(0)
        up = ctypes.c_int(0)
        _check(lib().yh_cg_mixed(self.ctx, ctypes.byref(par), d.h, s.h, ctypes.byref(it), ctypes.byref(up)))
        return it.value, up.value

    def cg_net(self, par, net, d, s):
        """Flexible CG on Hphi_sq preconditioned by net (None: plain CG
        through the same FP32 / FP64 reliable-update path)."""
        it = ctypes.c_int(0)
        up = ctypes.c_int(0)
        _check(lib().yh_cg_net(self.ctx, ctypes.byref(par), net.h if net is not None else None, d.h, s.h,
                               ctypes.byref(it), ctypes.byref(up)))
        return it.value, up.value

    def gcr_net(self, par, net, d, s, nkrylov=8):
        """Flexible GCR on Hphi_sq preconditioned by net (None: plain GCR),
        keeping the last nkrylov directions, FP32 inside FP64 reliable
        updates; returns (iterations, reliable updates)."""
        it = ctypes.c_int(0)
        up = ctypes.c_int(0)
        _check(lib().yh_gcr_net(self.ctx, ctypes.byref(par), net.h if net is not None else None, nkrylov, d.h, s.h,
                                ctypes.byref(it), ctypes.byref(up)))
        return it.value, up.value

    def cg_learned(self, par, net, mre, d, s):
        """The learned solve with its fallback (yh_cg_learned): flexible CG
        preconditioned by net from the MRE guess; plain mixed CG from the
        MRE guess if that does not converge.  Returns (iterations, reliable
        updates, fell_back)."""
        it, up, fb = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().yh_cg_learned(self.ctx, ctypes.byref(par), net.h if net is not None else None,
                                   mre.h if mre is not None else None, d.h, s.h, ctypes.byref(it), ctypes.byref(up),
                                   ctypes.byref(fb)))
        return it.value, up.value, bool(fb.value)

    # ---- forces (include/forces.h) ------------------------------------
    @property
    def ng(self):
        return self.nc ** 2 - 1

    def set_momenta(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        assert m.size == 4 * self.V * self.ng
        _check(lib().yh_set_momenta(self.ctx, _dp(m)))

    def get_momenta(self):
        out = np.zeros((4 * self.V, self.ng))
        _check(lib().yh_get_momenta(self.ctx, _dp(out)))
        return out

    def get_clover_force(self):
        out = np.zeros((6 * self.V, 2 * self.dr ** 2))
        _check(lib().yh_get_clover_force(self.ctx, _dp(out)))
        return out

    def begin_fermion_force(self):
        _check(lib().yh_begin_fermion_force(self.ctx))

    def end_fermion_force(self, dt):
        _check(lib().yh_end_fermion_force(self.ctx, dt))

    def clover_fermion_force(self, residue, lhs, rhs):
        _check(lib().yh_clover_fermion_force(self.ctx, residue, lhs.h, rhs.h))

    def wilson_fermion_force(self, dt, mass, residue, lhs, rhs):
        _check(lib().yh_wilson_fermion_force(self.ctx, dt, mass, residue, lhs.h, rhs.h))

    def wilson_fermion_force_q(self, dt, mass, residue, rhs):
        _check(lib().yh_wilson_fermion_force_q(self.ctx, dt, mass, residue, rhs.h))

    def clover_logdet_force(self, mass, residue):
        _check(lib().yh_clover_logdet_force(self.ctx, mass, residue))

    # ---- gauge molecular dynamics (forces.cpp:262-283, update.cpp:201-320) --
    def get_gauge(self):
        out = np.zeros((4 * self.V, 2 * self.nc ** 2))
        _check(lib().yh_get_gauge(self.ctx, _dp(out)))
        return out

    def represent_gauge_field(self):
        _check(lib().yh_represent_gauge_field(self.ctx))

    def wilson_gauge_force(self, dt, beta):
        _check(lib().yh_wilson_gauge_force(self.ctx, dt, beta))

    def improved_gauge_force(self, dt, beta, c0, c1):
        """improved_gauge_force (library/forces.cpp:285-310): plaquette
        staples times c0 plus the 1 x 2 rectangle staples times c1."""
        _check(lib().yh_improved_gauge_force(self.ctx, dt, beta, c0, c1))

    def update_links(self, eps):
        _check(lib().yh_update_links(self.ctx, eps))

    def reunitarize(self):
        _check(lib().yh_reunitarize(self.ctx))

    def plaquette(self):
        out = np.zeros(1)
        _check(lib().yh_plaquette(self.ctx, _dp(out)))
        return float(out[0])

    def avr_wilson(self, height, width):
        """avr_wilson(height, width) (library/wilsonloop.cpp:116-133)."""
        out = np.zeros(1)
        _check(lib().yh_avr_wilson(self.ctx, height, width, _dp(out)))
        return float(out[0])

    def momenta_energy(self):
        out = np.zeros(1)
        _check(lib().yh_momenta_energy(self.ctx, _dp(out)))
        return float(out[0])

    def apply_bcs_on_gauge_field(self):
        """apply_bcs_on_gauge_field (library/bcs.cpp:201-262) on the device
        fundamental links (no-op for periodic / antiperiodic)."""
        _check(lib().yh_apply_bcs_on_gauge_field(self.ctx))

    @property
    def gauge_volume(self):
        """The `volume` of the plaquette normalisation and the gauge action
        (library/bcs.cpp:32, 76; geometry.cpp:1150): (T-1) V3 for open,
        (T-2) V3 for SF boundaries, V otherwise."""
        v3 = self.V // self.dims[0]
        if self.bc == BC["open"]:
            return (self.dims[0] - 1) * v3
        if self.bc == BC["sf"]:
            return (self.dims[0] - 2) * v3
        return self.V

    # ---- measurement ---------------------------------------------------
    def bench_hop(self, prec, recon, reps, d, s):
        out = np.zeros(1)
        _check(lib().yh_bench_hop(self.ctx, PREC[prec] if isinstance(prec, str) else prec, recon, reps, d.h, s.h,
                                  _dp(out)))
        return out[0]

    def bench_copy(self, nbytes, reps):
        out = np.zeros(1)
        _check(lib().yh_bench_copy(self.ctx, int(nbytes), reps, _dp(out)))
        return out[0]

    def synchronize(self):
        _check(lib().yh_synchronize(self.ctx))

    def tune(self, key, value):
        _check(lib().yh_tune(self.ctx, key, value))

    def autotune(self, reps=20):
        """Pick the fastest XCD chunk per hop variant (yh_autotune); returns
        {(prec, recon): chunk} for the variants this context holds."""
        out = (ctypes.c_int * 6)()
        _check(lib().yh_autotune(self.ctx, reps, out))
        names = ("f64", "f32", "f16")
        return {(names[k // 2], 18 if k % 2 == 0 else 12): out[k] for k in range(6) if out[k] >= 0}


class Net:
    """The learned preconditioner: a gauge-covariant parallel-transport
    convolution network A on the even sublattice; M = A A^+."""

    def __init__(self, lat, channels=(1, 8, 8, 1)):
        self.lat = lat
        self.channels = tuple(int(c) for c in channels)
        ch = (ctypes.c_int * len(self.channels))(*self.channels)
        h = _VP()
        _check(lib().yh_net_create(lat.ctx, len(self.channels) - 1, ch, ctypes.byref(h)))
        self.h = h
        self.nw = lib().yh_net_num_weights(h)

    def shapes(self):
        c = self.channels
        return [(9, co, ci, 4, 4) for ci, co in zip(c[:-1], c[1:])]

    def _split(self, flat):
        out, k = [], 0
        for shp in self.shapes():
            n = int(np.prod(shp))
            out.append(flat[k:k + n].reshape(shp))
            k += n
        return out

    def set_weights(self, Ws):
        flat = np.concatenate([np.asarray(W, dtype=np.complex128).reshape(-1) for W in Ws])
        assert flat.size == self.nw
        buf = np.ascontiguousarray(flat).view(np.float64)
        _check(lib().yh_net_set_weights(self.h, _dp(buf)))

    def get_weights(self):
        buf = np.zeros(2 * self.nw)
        _check(lib().yh_net_get_weights(self.h, _dp(buf)))
        return self._split(buf.view(np.complex128))

    def init_weights(self, rng, scale=0.1, identity=0.5):
        """Small random weights plus `identity` on the path-0 channel diagonal."""
        Ws = []
        for shp in self.shapes():
            W = scale * (rng.normal(size=shp) + 1j * rng.normal(size=shp)) / np.sqrt(9 * shp[2] * 4)
            for k in range(min(shp[1], shp[2])):
                W[0, k, k] += identity * np.eye(4)
            Ws.append(W)
        self.set_weights(Ws)
        return Ws

    def apply(self, d, s, adjoint=False):
        _check(lib().yh_net_apply(self.h, int(adjoint), d.h, s.h))

    def precondition(self, d, s):
        _check(lib().yh_net_precondition(self.h, d.h, s.h))

    def loss_grad(self, bs, xs, want_grad=True):
        n = len(bs)
        B = (_VP * n)(*[b.h for b in bs])
        X = (_VP * n)(*[x.h for x in xs])
        loss = np.zeros(1)
        g = np.zeros(2 * self.nw) if want_grad else None
        _check(lib().yh_net_loss_grad(self.h, n, B, X, _dp(loss), _dp(g) if want_grad else None))
        return float(loss[0]), (self._split(g.view(np.complex128)) if want_grad else None)

    def adam_step(self, lr, npairs, beta1=0.9, beta2=0.999, eps=1e-8):
        _check(lib().yh_net_adam_step(self.h, lr, beta1, beta2, eps, npairs))

    def train(self, bs, xs, steps, batch=4, lr=2e-3, rng=None, log=None):
        """Adam on the device gradient of |A A^+ b - x|^2 / |x|^2 over
        minibatches of (source, solution) pairs; returns the loss history."""
        rng = rng or np.random.default_rng(0)
        hist = []
        for step in range(steps):
            idx = rng.choice(len(bs), size=min(batch, len(bs)), replace=False)
            loss, _ = self.loss_grad([bs[i] for i in idx], [xs[i] for i in idx], want_grad=False)
            self.adam_step(lr, len(idx))
            hist.append(loss / len(idx))
            if log and (step + 1) % log == 0:
                print(f"[net] step {step + 1} loss {hist[-1]:.4f}", flush=True)
        return hist

    def layer_host(self, layer, h, adjoint=False):
        """One layer on an activation (V, C, 4, 3) complex (testing)."""
        ci, co = self.channels[layer], self.channels[layer + 1]
        nout = ci if adjoint else co
        h = np.ascontiguousarray(h, dtype=np.complex128)
        out = np.zeros((self.lat.V, nout, 4, 3), dtype=np.complex128)
        _check(lib().yh_net_layer_host(self.h, layer, int(adjoint), _dp(h.view(np.float64)),
                                       _dp(out.view(np.float64))))
        return out

    def wgrad_host(self, layer, g, h):
        ci, co = self.channels[layer], self.channels[layer + 1]
        g = np.ascontiguousarray(g, dtype=np.complex128)
        h = np.ascontiguousarray(h, dtype=np.complex128)
        out = np.zeros((9, co, ci, 4, 4), dtype=np.complex128)
        _check(lib().yh_net_wgrad_host(self.h, layer, _dp(g.view(np.float64)), _dp(h.view(np.float64)),
                                       _dp(out.view(np.float64))))
        return out

    def free(self):
        if self.h:
            lib().yh_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MRE:
    """mre_par with mre_init / mre_guess / mre_store / mre_reset
    (library/mre.cpp, include/mre.h)."""

    def __init__(self, lat, max_past, mvm="Dphi", mass=0.0):
        self.lat = lat
        h = _VP()
        _check(lib().yh_mre_create(lat.ctx, max_past, MVM[mvm] if isinstance(mvm, str) else mvm, mass,
                                   ctypes.byref(h)))
        self.h = h

    def guess(self, d, s):
        _check(lib().yh_mre_guess(self.h, d.h, s.h))

    def store(self, x):
        _check(lib().yh_mre_store(self.h, x.h))

    def reset(self):
        _check(lib().yh_mre_reset(self.h))

    def free(self):
        if self.h:
            lib().yh_mre_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
