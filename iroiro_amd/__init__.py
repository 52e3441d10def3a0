"""iroiro_amd: the IroIro++ Wilson even/odd solve path on MI355X (gfx950).

Python mirror of the reference's operator interfaces for this path, backed
by the C ABI of libiroiro_hip.so (include/iroiro_hip.h):

  Dirac_Wilson_EvenOdd   lib/Dirac_ops/dirac_wilson_EvenOdd.hpp:12-92
  Dirac_Wilson           lib/Dirac_ops/dirac_wilson.hpp:19-175 (full index)
  Fopr_DdagD             lib/Fopr/fopr.h:146-153
  Solver_CG              lib/Solver/solver_CG.cpp:29-160
  MultiShiftSolver_CG    lib/Solver/multiShiftSolver_CG.cpp:10-185
  Inverter_WilsonLike    lib/Dirac_ops/eoUtils.cpp:43-80
  Qprop_EvenOdd          lib/Measurements/FermionicM/qprop_EvenOdd.cpp:11-43
  Qprop_MultiShift       lib/Measurements/FermionicM/qprop_MultiShift.cpp:10-31
  Source_local           lib/Measurements/FermionicM/source_types.hpp:23-62

Fields are float64 numpy arrays in the reference layouts (Format_F
spinors, Format_G links; even/odd half fields in esec/osec order).  All
arithmetic runs on the GPU; importing works without a GPU, but creating a
Context needs the HIP library and a device.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (ConvergenceErr, IIError, NotConverged, II_EVEN, II_ODD, II_FULL,
                   II_PREC_DOUBLE, II_PREC_SINGLE, II_PREC_HALF, II_PREC_BF16, PRECISIONS)

__all__ = [
    "Context", "Dirac_Wilson_EvenOdd", "Dirac_Wilson", "Fopr_DdagD", "Solver_CG",
    "MultiShiftSolver_CG", "MultiShiftSolver_Mixed", "Inverter_WilsonLike", "Qprop_EvenOdd", "Qprop_MultiShift",
    "Source_local",
    "SolverOutput", "DeviceField", "PropagatorSolver", "ConvergenceErr", "NotConverged",
    "IIError", "kappa_from_mass", "esec_osec", "LearnedPreconditioner", "Solver_FGCR",
]


def kappa_from_mass(mass):
    """kpp_ = 0.5/(4.0+mass) (lib/Dirac_ops/dirac_wilson.hpp:93)."""
    return 0.5 / (4.0 + mass)


def esec_osec(dims):
    """Lexicographic sites of the even and odd half lattices in ascending
    order (SiteIndex_EvenOdd::setup, lib/Geometry/siteIndex_EvenOdd.cpp:21-62)."""
    Lx, Ly, Lz, Lt = dims
    s = np.arange(Lx * Ly * Lz * Lt)
    x = s % Lx
    y = (s // Lx) % Ly
    z = (s // (Lx * Ly)) % Lz
    t = s // (Lx * Ly * Lz)
    par = (x + y + z + t) % 2
    return s[par == 0], s[par == 1]


class SolverOutput:
    """SolverOutput (lib/Solver/solver.hpp:16-22) plus device extras."""

    def __init__(self, c):
        self.diff = c.diff
        self.Iterations = c.iterations
        self.timing = c.timing_ms
        self.true_residual = c.true_residual
        self.restarts = c.restarts
        self.inner_iterations = c.inner_iterations
        self.recursion_restarts = c.recursion_restarts
        self.Msg = ""

    def __repr__(self):
        return ("SolverOutput(diff=%.3e, Iterations=%d, timing=%.3f ms, true_residual=%.3e, "
                "restarts=%d)" % (self.diff, self.Iterations, self.timing, self.true_residual,
                                  self.restarts))


class Context:
    """One lattice on one GPU: the device side of CommonPrms + Communicator
    for a single process (include/commonPrms.hpp:28-79)."""

    def __init__(self, dims, device=0):
        L = _lib.load()
        self.dims = tuple(int(d) for d in dims)
        if self.dims[0] % 2:
            raise ValueError("Lx must be even (lib/Geometry/siteIndex_EvenOdd.hpp:45)")
        h = ctypes.c_void_p()
        d = (ctypes.c_int * 4)(*self.dims)
        st = L.ii_context_create(ctypes.byref(h), d, int(device))
        if st != 0:
            raise IIError(st, "ii_context_create failed (is a GPU visible?)")
        self._h = h
        self.V = int(np.prod(self.dims))
        self.Vh = self.V // 2

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ii_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st, what):
        _lib.check(st, self._h, what)

    def upload_gauge(self, u):
        u = _lib.as_f64(u, 18 * 4 * self.V)
        self.check(_lib.load().ii_gauge_upload(self._h, _lib.dptr(u)), "ii_gauge_upload")

    def invalidate_gauge(self):
        self.check(_lib.load().ii_gauge_invalidate(self._h), "ii_gauge_invalidate")

    def download_gauge(self):
        u = np.empty(18 * 4 * self.V)
        self.check(_lib.load().ii_gauge_download(self._h, _lib.dptr(u)), "ii_gauge_download")
        return u

    def unit_gauge(self):
        """Cold start U = 1 on the device."""
        self.check(_lib.load().ii_gauge_unit(self._h), "ii_gauge_unit")

    def heatbath(self, beta, seed, nsweeps, first_sweep=0):
        """SU(3) heat-bath sweeps on the device gauge field."""
        self.check(_lib.load().ii_gauge_heatbath(self._h, float(beta), int(seed),
                                                 int(first_sweep), int(nsweeps)),
                   "ii_gauge_heatbath")

    def apply_apbc(self):
        """Antiperiodic time boundary on the device gauge field."""
        self.check(_lib.load().ii_gauge_apply_apbc(self._h), "ii_gauge_apply_apbc")

    def plaquette(self):
        p = ctypes.c_double()
        self.check(_lib.load().ii_gauge_plaquette(self._h, ctypes.byref(p)), "ii_gauge_plaquette")
        return p.value

    def device_info(self):
        buf = ctypes.create_string_buffer(256)
        g = ctypes.c_double()
        self.check(_lib.load().ii_device_info(self._h, buf, 256, ctypes.byref(g)), "ii_device_info")
        return buf.value.decode(), g.value

    def autotune(self):
        """Re-runs the launch-shape autotuner; returns its report."""
        buf = ctypes.create_string_buffer(8192)
        self.check(_lib.load().ii_autotune(self._h, buf, 8192), "ii_autotune")
        return buf.value.decode()

    def tune_save(self, path):
        """Writes the autotuner's choices of this context to a text file."""
        self.check(_lib.load().ii_tune_save(self._h, os.fsencode(path)), "ii_tune_save")

    def tune_load(self, path):
        """Loads autotuner choices saved on the same lattice (no timing runs)."""
        self.check(_lib.load().ii_tune_load(self._h, os.fsencode(path)), "ii_tune_load")

    def bench_copy(self, nbytes, nrep=20):
        """Average ms of a plain device copy of nbytes (read + write)."""
        ms = ctypes.c_double()
        self.check(_lib.load().ii_bench_copy(self._h, int(nbytes), int(nrep), ctypes.byref(ms)),
                   "ii_bench_copy")
        return ms.value

    def read_gauge(self, path, fmt="NERSC", check=False):
        """GaugeGlobal::initialize with Type Binary / NERSC / ILDG
        (lib/Main/gaugeGlobal.cpp:133-180): the file goes straight into the
        device gauge field; returns the ii_io_info fields as a dict."""
        info = _lib.IOInfoC()
        self.check(_lib.load().ii_gauge_read(self._h, os.fsencode(path), _lib.II_IO[fmt],
                                             1 if check else 0, ctypes.byref(info)),
                   "ii_gauge_read")
        return {k: getattr(info, k) for k, _ in info._fields_}

    def write_gauge(self, path, fmt="NERSC"):
        """CCIO::SaveOnDisk<Format_G> with the Binary / NERSC / ILDG writer."""
        info = _lib.IOInfoC()
        self.check(_lib.load().ii_gauge_write(self._h, os.fsencode(path), _lib.II_IO[fmt],
                                              ctypes.byref(info)), "ii_gauge_write")
        return {k: getattr(info, k) for k, _ in info._fields_}

    def bench_md_force(self, nrep=20):
        """(force kernel ms, md_force ms incl. its two hops) on device fields."""
        f, t = ctypes.c_double(), ctypes.c_double()
        self.check(_lib.load().ii_bench_md_force(self._h, nrep, ctypes.byref(f),
                                                 ctypes.byref(t)), "ii_bench_md_force")
        return f.value, t.value

    def bench_hop(self, precision="double", nrhs=1, nrep=50):
        ms = ctypes.c_double()
        self.check(_lib.load().ii_bench_hop(self._h, PRECISIONS[precision], nrhs, nrep,
                                            ctypes.byref(ms)), "ii_bench_hop")
        return ms.value


class _OpBase:
    def __init__(self, mass, u, ctx):
        self.ctx = ctx
        self.mass = float(mass)
        self.kpp = kappa_from_mass(self.mass)
        self.u = u
        self.update_internal_state()

    def getKappa(self):
        return self.kpp

    def getMass(self):
        return 0.5 / self.kpp - 4.0

    def getGaugeField_ptr(self):
        return self.u

    def gsize(self):
        return 18 * 4 * self.ctx.V

    def update_internal_state(self):
        """Dirac::update_internal_state (lib/Dirac_ops/dirac.hpp:28): the
        caller changed the gauge field; upload it again."""
        self.ctx.upload_gauge(self.u)

    def _call(self, name, f, nout, *extra):
        f = _lib.as_f64(f)
        out = np.empty(nout)
        fn = getattr(_lib.load(), name)
        self.ctx.check(fn(self.ctx.handle, self.kpp, _lib.dptr(out), _lib.dptr(f), *extra), name)
        return out


class Dirac_Wilson_EvenOdd(_OpBase):
    """Dirac_Wilson_EvenOdd<Dirac_Wilson> (dirac_wilson_EvenOdd.hpp:12-92):
    the Schur operator Dhat = 1 - Deo Doe on even sites."""

    def __init__(self, mass, u, ctx):
        super().__init__(mass, u, ctx)

    def fsize(self):
        return 24 * self.ctx.Vh

    def mult(self, f):
        return self._call("ii_dirac_mult", _lib.as_f64(f, self.fsize()), self.fsize())

    def mult_dag(self, f):
        return self._call("ii_dirac_mult_dag", _lib.as_f64(f, self.fsize()), self.fsize())

    def mult_eo(self, f):
        """Deo.mult: odd -> even (:59-62)."""
        return self._call("ii_dirac_mult_eo", _lib.as_f64(f, self.fsize()), self.fsize())

    def mult_oe(self, f):
        """Doe.mult: even -> odd (:64-67)."""
        return self._call("ii_dirac_mult_oe", _lib.as_f64(f, self.fsize()), self.fsize())

    def mult_eo_dag(self, f):
        """Doe.mult_dag: even -> odd (:69-72)."""
        return self._call("ii_dirac_mult_eo_dag", _lib.as_f64(f, self.fsize()), self.fsize())

    def mult_oe_dag(self, f):
        """Deo.mult_dag: odd -> even (:74-77)."""
        return self._call("ii_dirac_mult_oe_dag", _lib.as_f64(f, self.fsize()), self.fsize())

    def mult_eo_prec(self, f, precision, out_par=0, dag=False):
        """The hop to parity out_par (mult_eo / mult_oe, or mult_oe_dag /
        mult_eo_dag with dag) in a storage precision ("single", "half",
        "bf16", "double"), as the mixed-precision solvers apply it.  f holds
        nrhs half fields back to back; nrhs > 1 runs the fused multi-rhs
        kernel.  Returns FP64."""
        n = self.fsize()
        f = np.ascontiguousarray(f, dtype=np.float64).ravel()
        if f.size % n:
            raise ValueError("mult_eo_prec: size %d is not a multiple of %d" % (f.size, n))
        out = np.empty_like(f)
        self.ctx.check(_lib.load().ii_dirac_mult_eo_prec(
            self.ctx.handle, self.kpp, PRECISIONS[precision], int(out_par), int(bool(dag)),
            f.size // n, _lib.dptr(out), _lib.dptr(f)), "ii_dirac_mult_eo_prec")
        return out

    def mult_ee(self, f):
        return np.array(f, dtype=np.float64)

    mult_oo = mult_ee_inv = mult_oo_inv = mult_ee

    def md_force(self, eta, zeta):
        """Dirac_Wilson_EvenOdd::md_force (:110-116) from even half fields:
        a gauge-layout force field (18*4*V doubles)."""
        eta = _lib.as_f64(eta, self.fsize())
        zeta = _lib.as_f64(zeta, self.fsize())
        out = np.empty(self.gsize())
        self.ctx.check(_lib.load().ii_dirac_md_force(self.ctx.handle, self.kpp, _lib.dptr(out),
                                                     _lib.dptr(eta), _lib.dptr(zeta)),
                       "ii_dirac_md_force")
        return out

    def gamma5(self, f):
        f = _lib.as_f64(f)
        out = np.empty_like(f)
        self.ctx.check(_lib.load().ii_dirac_gamma5(self.ctx.handle, _lib.dptr(out), _lib.dptr(f),
                                                   f.size // 24), "ii_dirac_gamma5")
        return out


class Dirac_Wilson(_OpBase):
    """Full-index Dirac_Wilson (dirac_wilson.hpp:118-134): D = 1 - kappa H."""

    def fsize(self):
        return 24 * self.ctx.V

    def mult(self, f):
        return self._call("ii_dirac_full_mult", _lib.as_f64(f, self.fsize()), self.fsize(), 0)

    def mult_dag(self, f):
        return self._call("ii_dirac_full_mult", _lib.as_f64(f, self.fsize()), self.fsize(), 1)

    def md_force(self, eta, zeta):
        """Dirac_Wilson::md_force (dirac_wilson.cpp:52-60): -kappa (p + m)."""
        eta = _lib.as_f64(eta, self.fsize())
        zeta = _lib.as_f64(zeta, self.fsize())
        out = np.empty(self.gsize())
        self.ctx.check(_lib.load().ii_dirac_full_md_force(self.ctx.handle, self.kpp,
                                                          _lib.dptr(out), _lib.dptr(eta),
                                                          _lib.dptr(zeta)),
                       "ii_dirac_full_md_force")
        return out

    def _md_force_pm(self, fce, eta, zeta, minus):
        if fce.dtype != np.float64 or not fce.flags.c_contiguous or fce.size != self.gsize():
            raise ValueError("fce must be a contiguous float64 array of gsize() doubles")
        eta = _lib.as_f64(eta, self.fsize())
        zeta = _lib.as_f64(zeta, self.fsize())
        self.ctx.check(_lib.load().ii_dirac_full_md_force_pm(self.ctx.handle, _lib.dptr(fce),
                                                             _lib.dptr(eta), _lib.dptr(zeta),
                                                             minus),
                       "ii_dirac_full_md_force_pm")

    def md_force_p(self, fce, eta, zeta):
        """Dirac_Wilson::md_force_p (dirac_wilson.cpp:41-43): adds into fce."""
        self._md_force_pm(fce, eta, zeta, 0)

    def md_force_m(self, fce, eta, zeta):
        """Dirac_Wilson::md_force_m (dirac_wilson.cpp:45-50): adds into fce."""
        self._md_force_pm(fce, eta, zeta, 1)

    def gamma5(self, f):
        f = _lib.as_f64(f)
        out = np.empty_like(f)
        self.ctx.check(_lib.load().ii_dirac_gamma5(self.ctx.handle, _lib.dptr(out), _lib.dptr(f),
                                                   f.size // 24), "ii_dirac_gamma5")
        return out


class Fopr_DdagD:
    """Fopr_DdagD (lib/Fopr/fopr.h:146-153): mult = D^dag D, Hermitian."""

    def __init__(self, D):
        assert D is not None
        self.D = D

    def func(self, x):
        return x

    def fsize(self):
        return self.D.fsize()

    def mult(self, f):
        if isinstance(self.D, Dirac_Wilson_EvenOdd):
            return self.D._call("ii_fopr_ddagd_mult", _lib.as_f64(f, self.fsize()), self.fsize())
        return self.D.mult_dag(self.D.mult(f))

    def mult_dag(self, f):
        return self.mult(f)


class Solver_CG:
    """Solver_CG (lib/Solver/solver_CG.cpp:29-160), the whole iteration on
    the device.  Stops when |r|^2/|b|^2 < prec (note: squared), starting from
    x0 = b.  Raises ConvergenceErr where the reference exits."""

    def __init__(self, prec, maxiter, fopr):
        self.prec = float(prec)
        self.MaxIter = int(maxiter)
        self.opr = fopr

    def check_DdagD(self):
        return isinstance(self.opr, Fopr_DdagD)

    def solve(self, sol, src):
        D = self.opr.D
        if not isinstance(D, Dirac_Wilson_EvenOdd):
            raise TypeError("Solver_CG on the device supports Fopr_DdagD(Dirac_Wilson_EvenOdd)")
        b = _lib.as_f64(src, D.fsize())
        x = np.empty(D.fsize())
        out = _lib.SolverOutputC()
        st = _lib.load().ii_solver_cg(D.ctx.handle, D.kpp, _lib.dptr(x), _lib.dptr(b), self.prec,
                                      self.MaxIter, ctypes.byref(out))
        sol[...] = x
        D.ctx.check(st, "Solver_CG")
        so = SolverOutput(out)
        so.Msg = "CG solver"
        return so


def TracelessAntihermite(ctx, G):
    """FieldUtils::TracelessAntihermite (lib/Tools/fieldUtils.cpp:264-274)
    on a gauge-layout field: (m - m^dag)/2 - i Im tr(m)/3 per link."""
    G = _lib.as_f64(G, 18 * 4 * ctx.V)
    out = np.empty_like(G)
    ctx.check(_lib.load().ii_traceless_antihermite(ctx.handle, _lib.dptr(out), _lib.dptr(G)),
              "ii_traceless_antihermite")
    return out


class Action_Nf2:
    """Action_Nf2 (lib/Action/action_Nf2.cpp:14-86, no smearing) with an
    even/odd Wilson operator and a Solver_CG: the two-flavour pseudofermion
    action.  md_force runs the solve, Dhat eta, the MD force and the
    traceless anti-hermitian projection on the device in one call."""

    def __init__(self, D, solver):
        if not isinstance(D, Dirac_Wilson_EvenOdd) or not isinstance(solver, Solver_CG):
            raise TypeError("Action_Nf2 on the device takes Dirac_Wilson_EvenOdd and Solver_CG")
        self.D = D
        self.slv = solver
        self.phi = None

    def init(self, seed):
        """phi = D^dag xi with Gaussian xi (action_Nf2.cpp:41-53); the
        reference's MPrand generator is replaced by numpy's, seeded."""
        xi = np.random.default_rng(seed).standard_normal(self.D.fsize())
        self.phi = self.D.mult_dag(xi)

    def set_phi(self, phi):
        self.phi = _lib.as_f64(phi, self.D.fsize()).copy()

    def DdagD_inv(self, src):
        sol = np.empty(self.D.fsize())
        self.slv.solve(sol, src)
        return sol

    def calc_H(self):
        """phi . (D^dag D)^-1 phi (action_Nf2.cpp:55-67)."""
        return float(np.dot(self.phi, self.DdagD_inv(self.phi)))

    def md_force(self):
        """Action_Nf2::md_force (action_Nf2.cpp:69-86)."""
        out = _lib.SolverOutputC()
        force = np.empty(self.D.gsize())
        st = _lib.load().ii_action_nf2_md_force(self.D.ctx.handle, self.D.kpp,
                                                _lib.dptr(self.phi), self.slv.prec,
                                                self.slv.MaxIter, _lib.dptr(force),
                                                ctypes.byref(out))
        self.D.ctx.check(st, "Action_Nf2::md_force")
        self.last_solve = SolverOutput(out)
        return force


class MultiShiftSolver_CG:
    """MultiShiftSolver_CG (lib/Solver/multiShiftSolver_CG.cpp:10-185)."""

    def __init__(self, fopr, stop_cond, niter):
        self.opr = fopr
        self.prec = float(stop_cond)
        self.MaxIter = int(niter)

    def solve(self, solution, source, shifts):
        """Fills solution[i] for every shift; returns (SolverOutput, diff,
        Nconv) with the reference's meanings (diff = max squared residual,
        SolverOutput.diff = its square root)."""
        D = self.opr.D
        b = _lib.as_f64(source, D.fsize())
        sh = _lib.as_f64(shifts)
        x = np.empty(len(sh) * D.fsize())
        out = _lib.SolverOutputC()
        # the even/odd Schur operator or the full-index one (Qprop_MultiShift)
        fn = (_lib.load().ii_multishift_cg if isinstance(D, Dirac_Wilson_EvenOdd)
              else _lib.load().ii_multishift_cg_full)
        st = fn(D.ctx.handle, D.kpp, _lib.dptr(x), _lib.dptr(b), _lib.dptr(sh), len(sh),
                self.prec, self.MaxIter, ctypes.byref(out))
        xs = x.reshape(len(sh), D.fsize())
        for i in range(len(sh)):
            solution[i][...] = xs[i]
        if st == _lib.II_ERR_NOT_CONVERGED:
            raise NotConverged(st, "Not converged.")
        D.ctx.check(st, "MultiShiftSolver_CG")
        so = SolverOutput(out)
        so.Msg = "Multishift CG solver"
        so.diff = out.true_residual
        return so, out.diff, out.iterations


class RationalApprox:
    """RationalApprox (lib/Tools/RationalApprox/rationalapprox.cpp:11-103,
    rationalapprox.hpp:56-96): the minimax relative-error approximations
        x^(a/b)  ~ Const()    + sum Residuals()[i]    / (x + Poles()[i])
        x^(-a/b) ~ InvConst() + sum InvResiduals()[i] / (x + InvPoles()[i])
    on [low, high], computed by the library's quad-precision Remez
    (ii_rational_approx).  `precision` (the reference's MPFR digits) is
    accepted for the same parameter list; the arithmetic is 113-bit."""

    def __init__(self, num_deg, den_deg, exp_num, exp_den, precision, low, high):
        if num_deg != den_deg:
            raise ValueError("numerator and denominator degrees must be equal "
                             "(rationalapprox.cpp:45)")
        self.params = dict(num_deg=int(num_deg), den_deg=int(den_deg), exp_num=int(exp_num),
                           exp_den=int(exp_den), precision=int(precision), low=float(low),
                           high=float(high))
        n = int(num_deg)
        res, pole, ires, ipole = (np.empty(n) for _ in range(4))
        a0, ia0, err = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = _lib.load().ii_rational_approx(None, n, int(den_deg), int(exp_num), int(exp_den),
                                            float(low), float(high), _lib.dptr(res),
                                            _lib.dptr(pole), ctypes.byref(a0), _lib.dptr(ires),
                                            _lib.dptr(ipole), ctypes.byref(ia0),
                                            ctypes.byref(err))
        if st != _lib.II_OK:
            raise IIError(st, "ii_rational_approx: the Remez exchange failed for degree %d on "
                              "[%g, %g]" % (n, low, high))
        self.RA_res, self.RA_pole, self.RA_a0 = res, pole, a0.value
        self.I_RA_res, self.I_RA_pole, self.I_RA_a0 = ires, ipole, ia0.value
        self.error = err.value

    def exponent(self):
        return self.params["exp_num"] / self.params["exp_den"]

    def Residuals(self):
        return self.RA_res.copy()

    def Poles(self):
        return self.RA_pole.copy()

    def Const(self):
        return self.RA_a0

    def InvResiduals(self):
        return self.I_RA_res.copy()

    def InvPoles(self):
        return self.I_RA_pole.copy()

    def InvConst(self):
        return self.I_RA_a0

    def rescale(self, min_in, max_in):
        """rationalapprox.cpp:82-103: the coefficients are taken as computed
        on [a, 1] and stretched to 1.05 * max (min enters only as 0.95 min,
        unused by the reference)."""
        mx = 1.05 * max_in
        factor = mx ** self.exponent()
        self.RA_a0 *= factor
        self.RA_res = self.RA_res * (factor * mx)
        self.RA_pole = self.RA_pole * mx
        factor = mx ** (-self.exponent())
        self.I_RA_a0 *= factor
        self.I_RA_res = self.I_RA_res * (factor * mx)
        self.I_RA_pole = self.I_RA_pole * mx


class RationalSolver_CG:
    """RationalSolver_CG (lib/Solver/rationalSolver_CG.cpp:12-111) over the
    device multishift of a MultiShiftSolver_CG on Fopr_DdagD(Dirac_Wilson_
    EvenOdd); the reconstruction sum runs on the device too.  precision
    "double" is the reference's FP64 recursion (stop_cond = its squared
    test); "single"/"half"/"bf16" use the mixed-precision multishift with
    stop_cond as the target true residual of every shift."""

    def __init__(self, ms_solver, approx=None, precision="double"):
        self.MS_Solver_ = ms_solver
        self.precision = PRECISIONS[precision]
        self.Residuals = self.Poles = self.InvResiduals = self.InvPoles = None
        self.ConstTerm = self.InvConstTerm = 0.0
        if approx is not None:
            self.set_Approx(approx)

    def set_Approx(self, RA):
        self.Residuals, self.Poles, self.ConstTerm = RA.Residuals(), RA.Poles(), RA.Const()
        self.InvResiduals, self.InvPoles = RA.InvResiduals(), RA.InvPoles()
        self.InvConstTerm = RA.InvConst()

    def check_DdagD(self):
        return True

    def _D(self):
        D = self.MS_Solver_.opr.D
        if not isinstance(D, Dirac_Wilson_EvenOdd):
            raise TypeError("RationalSolver_CG on the device runs on Dirac_Wilson_EvenOdd")
        return D

    def _solve(self, sol, src, a0, res, poles, what):
        if res is None or len(res) == 0:
            raise RuntimeError("[RationalSolver] Rational Approximation not initialized yet")
        D = self._D()
        b = _lib.as_f64(src, D.fsize())
        x = np.empty(D.fsize())
        res = _lib.as_f64(res)
        poles = _lib.as_f64(poles)
        out = _lib.SolverOutputC()
        st = _lib.load().ii_rational_solve(D.ctx.handle, D.kpp, _lib.dptr(x), _lib.dptr(b), a0,
                                           _lib.dptr(res), _lib.dptr(poles), len(res),
                                           self.MS_Solver_.prec, self.MS_Solver_.MaxIter,
                                           self.precision, ctypes.byref(out))
        sol[...] = x
        if st == _lib.II_ERR_NOT_CONVERGED:
            raise NotConverged(st, "Not converged.")
        D.ctx.check(st, what)
        so = SolverOutput(out)
        so.Msg = what
        return so

    def solve(self, sol, src):
        """(M^dag M)^(a/b) src."""
        return self._solve(sol, src, self.ConstTerm, self.Residuals, self.Poles,
                           "RationalSolver_CG::solve")

    def solve_inv(self, sol, src):
        """(M^dag M)^(-a/b) src."""
        return self._solve(sol, src, self.InvConstTerm, self.InvResiduals, self.InvPoles,
                           "RationalSolver_CG::solve_inv")

    def _no_reconstruct(self, shifted_sol, src, poles):
        if poles is None or len(poles) == 0:
            raise RuntimeError("[RationalSolver] Rational Approximation not initialized yet")
        shifted_sol[:] = [np.empty(self._D().fsize()) for _ in range(len(poles))]
        so, _, _ = self.MS_Solver_.solve(shifted_sol, src, poles)
        return so

    def solve_noReconstruct(self, shifted_sol, src):
        """The shifted solutions with the inverse poles (force term)."""
        return self._no_reconstruct(shifted_sol, src, self.InvPoles)

    def solve_noReconstruct_inv(self, shifted_sol, src):
        return self._no_reconstruct(shifted_sol, src, self.Poles)


class Action_Nf:
    """Action_Nf (lib/Action/action_Nf.cpp, action_Nf.hpp:61-140), the
    RHMC term for n_flav flavours with n_pseudof pseudofermions on an
    even/odd Wilson kernel, no smearing.  Approximations, as the reference
    builds them: Metropolis x^(Nf/2n), molecular dynamics x^(Nf/2n) (used
    through its inverse), pseudofermions x^(Nf/4n); degree, precision and
    bounds per step in the order Metropolis, MD, PF.  md_force runs the
    shifted solves, the per-pole MD force and the projection on the device
    in one call."""

    def __init__(self, D, slv, n_flav, n_pseudof, degree, precision, b_low, b_high):
        self.D, self.slv = D, slv
        self.n_flav, self.n_pseudof = int(n_flav), int(n_pseudof)
        mk = lambda k, den: RationalApprox(degree[k], degree[k], self.n_flav, den,
                                           precision[k], b_low[k], b_high[k])
        self.MetropolisApprox_ = mk(0, 2 * self.n_pseudof)
        self.MolecularDynApprox_ = mk(1, 2 * self.n_pseudof)
        self.PseudoFermionsApprox_ = mk(2, 4 * self.n_pseudof)
        self.phi_ = [np.zeros(D.fsize()) for _ in range(self.n_pseudof)]

    def init(self, seed):
        """phi_i = (M^dag M)^(Nf/4n) xi_i (action_Nf.cpp:39-57); numpy's
        Gaussian generator, seeded, in place of the reference's MPrand."""
        self.slv.set_Approx(self.PseudoFermionsApprox_)
        rng = np.random.default_rng(seed)
        for i in range(self.n_pseudof):
            self.slv.solve(self.phi_[i], rng.standard_normal(self.D.fsize()))

    def calc_H(self):
        """sum_i phi_i . (M^dag M)^(-Nf/2n) phi_i (action_Nf.cpp:59-79)."""
        self.slv.set_Approx(self.MetropolisApprox_)
        H = 0.0
        tmp = np.empty(self.D.fsize())
        for i in range(self.n_pseudof):
            self.slv.solve_inv(tmp, self.phi_[i])
            H += float(np.dot(self.phi_[i], tmp))
        return H

    def md_force(self):
        """action_Nf.cpp:81-110 on the device (ii_action_nf_md_force)."""
        self.slv.set_Approx(self.MolecularDynApprox_)
        ms = self.slv.MS_Solver_
        phi = np.concatenate(self.phi_)
        ires = _lib.as_f64(self.MolecularDynApprox_.InvResiduals())
        ipoles = _lib.as_f64(self.MolecularDynApprox_.InvPoles())
        force = np.empty(self.D.gsize())
        out = _lib.SolverOutputC()
        st = _lib.load().ii_action_nf_md_force(self.D.ctx.handle, self.D.kpp, _lib.dptr(phi),
                                               self.n_pseudof, _lib.dptr(ires),
                                               _lib.dptr(ipoles), len(ires), ms.prec,
                                               ms.MaxIter, self.slv.precision,
                                               _lib.dptr(force), ctypes.byref(out))
        if st == _lib.II_ERR_NOT_CONVERGED:
            raise NotConverged(st, "Not converged.")
        self.D.ctx.check(st, "Action_Nf::md_force")
        self.last_solve = SolverOutput(out)
        return force


MESON_TYPES = ["Pion", "Scalar", "Vector1", "Vector2", "Vector3", "Vector4", "AVector1",
               "AVector2", "AVector3", "AVector4", "Tensor12", "Tensor13", "Tensor14",
               "Tensor23", "Tensor24", "Tensor34", "AV4_PS", "PS_AV4"]


def _prop_field(ctx, q):
    """A 12-column propagator as a device field: a DeviceField is used as
    is; a list of 12 host full fields (prop_t) is uploaded."""
    if isinstance(q, DeviceField):
        if q.subset != "full" or q.nrhs != 12:
            raise ValueError("a propagator is a full DeviceField of 12 columns")
        return q
    f = DeviceField(ctx, "full", 12)
    for k in range(12):
        f.upload(k, q[k])
    return f


def meson_correlators(ctx, q1, q2, mode=0, direction=3, mom=(0.0, 0.0, 0.0)):
    """All 18 MesonTypes at once (ii_meson_correlators): mode 0 calculate,
    1 calculate_mom, 2 calculate_PSmom; returns an array [18][...]."""
    f1 = _prop_field(ctx, q1)
    f2 = f1 if q2 is q1 else _prop_field(ctx, q2)
    L = ctx.dims[direction]
    n = 2 * L if mode == 1 else L
    out = np.empty(18 * n)
    m = _lib.as_f64(mom, 3)
    ctx.check(_lib.load().ii_meson_correlators(ctx.handle, f1.handle, f2.handle, int(mode),
                                               int(direction), _lib.dptr(m), _lib.dptr(out)),
              "ii_meson_correlators")
    return out.reshape(18, n)


class MesonCorrelator:
    """MesonCorrelator (lib/Measurements/FermionicM/meson_correlator.hpp:
    26-136) on the device: q1, q2 are 12-column propagators (DeviceField or
    a list of 12 host full fields)."""

    def __init__(self, ctx, mtype="Pion"):
        self.ctx = ctx
        self.type = MESON_TYPES.index(mtype) if isinstance(mtype, str) else int(mtype)

    def calculate(self, q1, q2, direction=3):
        return meson_correlators(self.ctx, q1, q2, 0, direction)[self.type]

    def calculate_mom(self, q1, q2, a, b, c):
        return meson_correlators(self.ctx, q1, q2, 1, 3, (a, b, c))[self.type]

    def calculate_PSmom(self, q1, q2, a, b, c):
        return meson_correlators(self.ctx, q1, q2, 2, 3, (a, b, c))[self.type]


class MultiShiftSolver_Mixed:
    """Mixed-precision multishift (no reference counterpart): all shifts
    fused in an FP32 shifted recursion, then a batched reliable-update CG
    per shift (inner precision `precision`, FP64 true residuals) until
    |b - (A + sigma_i) x_i| <= tol |b| for every shift."""

    def __init__(self, fopr, tol=1e-10, precision="single", maxiter=100000):
        self.opr = fopr
        self.tol = float(tol)
        self.precision = PRECISIONS[precision]
        self.maxiter = int(maxiter)

    def solve(self, solution, source, shifts):
        D = self.opr.D
        b = _lib.as_f64(source, D.fsize())
        sh = _lib.as_f64(shifts)
        x = np.empty(len(sh) * D.fsize())
        outs = (_lib.SolverOutputC * len(sh))()
        st = _lib.load().ii_multishift_mixed(D.ctx.handle, D.kpp, _lib.dptr(x), _lib.dptr(b),
                                             _lib.dptr(sh), len(sh), self.tol, self.precision,
                                             self.maxiter, outs)
        xs = x.reshape(len(sh), D.fsize())
        for i in range(len(sh)):
            solution[i][...] = xs[i]
        D.ctx.check(st, "MultiShiftSolver_Mixed")
        return [SolverOutput(o) for o in outs]


class Inverter_WilsonLike:
    """EvenOddUtils::Inverter_WilsonLike (lib/Dirac_ops/eoUtils.cpp:43-80)."""

    def __init__(self, D, solver):
        self.D = D
        self.slv = solver

    def invert(self, sol, src):
        D = self.D
        b = _lib.as_f64(src, 24 * D.ctx.V)
        x = np.empty(24 * D.ctx.V)
        out = _lib.SolverOutputC()
        st = _lib.load().ii_eo_invert(D.ctx.handle, D.kpp, _lib.dptr(x), _lib.dptr(b),
                                      self.slv.prec, self.slv.MaxIter, ctypes.byref(out))
        sol[...] = x
        D.ctx.check(st, "Inverter_WilsonLike::invert")
        return SolverOutput(out)


class Source_local:
    """Point source (source_types.hpp:23-62): 1.0 in the real part of
    (colour c, spin s) at one site."""

    def __init__(self, pos, dims):
        self.pos = tuple(int(p) for p in pos)
        self.dims = tuple(dims)
        Lx, Ly, Lz, _ = self.dims
        x, y, z, t = self.pos
        self.loc = x + Lx * (y + Ly * (z + Lz * t))
        self.V = int(np.prod(self.dims))

    def mksrc(self, s, c):
        f = np.zeros(24 * self.V)
        f[24 * self.loc + 2 * (c + 3 * s)] = 1.0
        return f


class Qprop_EvenOdd:
    """Qprop_EvenOdd::calc (qprop_EvenOdd.cpp:11-23): 12 inversions, column
    3*s + c."""

    def __init__(self, D, solver):
        self.inv = Inverter_WilsonLike(D, solver)

    def calc(self, src):
        xq = []
        for s in range(4):
            for c in range(3):
                sol = np.empty(24 * self.inv.D.ctx.V)
                self.inv.invert(sol, src.mksrc(s, c))
                xq.append(sol)
        return xq


class Qprop_MultiShift:
    """Qprop_MultiShift::calc (qprop_MultiShift.cpp:10-31): for each column
    3*s + c a multishift solve of (D^dag D + sigma_m) x_m = src on the
    full-index operator, then xq[(3s+c)*nshift + m] = D x_m.  Runs on the
    device (one call for all 12 columns)."""

    def __init__(self, D, solver):
        if not isinstance(D, Dirac_Wilson):
            raise TypeError("Qprop_MultiShift takes the full-index Dirac_Wilson")
        self.D = D
        self.slv = solver

    def calc(self, source, mass_shifts):
        ctx = self.D.ctx
        sh = _lib.as_f64(mass_shifts)
        n = len(sh)
        xqf = DeviceField(ctx, "full", 12 * n)
        outs = (_lib.SolverOutputC * 12)()
        site = (ctypes.c_int * 4)(*source.pos)
        try:
            st = _lib.load().ii_qprop_multishift(ctx.handle, self.D.kpp, site, _lib.dptr(sh), n,
                                                 self.slv.prec, self.slv.MaxIter, xqf.handle,
                                                 outs)
            if st == _lib.II_ERR_NOT_CONVERGED:
                raise NotConverged(st, "Not converged.")
            ctx.check(st, "Qprop_MultiShift")
            xq = [xqf.download(k) for k in range(12 * n)]
        finally:
            xqf.close()
        self.outputs = [SolverOutput(o) for o in outs]
        return xq


class DeviceField:
    """A device-resident field (ii_field): nrhs right-hand sides of one
    subset ('even', 'odd' or 'full')."""

    _SUB = {"even": II_EVEN, "odd": II_ODD, "full": II_FULL}

    def __init__(self, ctx, subset="full", nrhs=1):
        self.ctx = ctx
        self.subset = subset
        self.nrhs = int(nrhs)
        h = ctypes.c_void_p()
        ctx.check(_lib.load().ii_field_create(ctx.handle, self._SUB[subset], self.nrhs,
                                              ctypes.byref(h)), "ii_field_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def size(self):
        return 24 * (self.ctx.V if self.subset == "full" else self.ctx.Vh)

    def upload(self, rhs, host):
        a = _lib.as_f64(host, self.size)
        self.ctx.check(_lib.load().ii_field_upload(self._h, int(rhs), _lib.dptr(a)),
                       "ii_field_upload")

    def download(self, rhs):
        a = np.empty(self.size)
        self.ctx.check(_lib.load().ii_field_download(self._h, int(rhs), _lib.dptr(a)),
                       "ii_field_download")
        return a

    def gaussian(self, rhs, seed, stream):
        """Complex Gaussian source on the even sites of rhs (odd sites zero)."""
        self.ctx.check(_lib.load().ii_field_gaussian(self.handle, int(rhs), int(seed),
                                                     int(stream)), "ii_field_gaussian")

    def point_sources(self, site):
        d = (ctypes.c_int * 4)(*site)
        self.ctx.check(_lib.load().ii_field_point_sources(self._h, d), "ii_field_point_sources")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().ii_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PropagatorSolver:
    """The production solve of many right-hand sides (e.g. the 12 columns of
    a point-source propagator) through the even/odd Schur system, entirely
    on the device, stopping on the FP64 true residual |b - D x|/|b| per rhs.

    precision: 'double' | 'single' | 'half' | 'bf16' (inner storage);
    batched: all rhs in one fused multi-rhs solve;
    method: 'cg' (CG on Dhat^dag Dhat, the reference's Solver_CG on
    Fopr_DdagD) or 'bicgstab' (BiCGStab on Dhat, sloppy precisions only)."""

    def __init__(self, ctx, kappa, tol=1e-10, precision="double", batched=False, maxiter=100000,
                 delta=0.0, use_graph=False, use_learned=False, method="cg",
                 restart_factor=0.0, nkrylov=0, northo=0):
        self.ctx = ctx
        self.kappa = float(kappa)
        self.params = _lib.PropParamsC(tol, maxiter, PRECISIONS[precision], int(batched),
                                       float(delta), int(use_graph), int(use_learned),
                                       _lib.II_METHOD[method], float(restart_factor),
                                       int(nkrylov), int(northo))

    def solve(self, src, sol):
        outs = (_lib.SolverOutputC * src.nrhs)()
        st = _lib.load().ii_propagator_solve(self.ctx.handle, self.kappa, ctypes.byref(self.params),
                                             src.handle, sol.handle, outs)
        res = [SolverOutput(o) for o in outs]
        self.ctx.check(st, "ii_propagator_solve")
        return res


class LearnedPreconditioner:
    """The trained gauge-covariant preconditioner M of the even/odd Schur
    operator (no reference counterpart): a linear network of 4 layers of
    widths 1 -> 8 -> 8 -> 8 -> 1 spinor channels, each
      out[o](x) = sum_d sum_i W[d][o][i] (T_d in[i])(x)
    over the 9 nearest-neighbour parallel-transport paths d, W complex 4x4
    spin matrices.  Weights: complex numpy arrays [9][cout][cin][4][4]."""

    WIDTHS = _lib.NET_WIDTHS

    def __init__(self, ctx):
        self.ctx = ctx

    def _L(self):
        return _lib.load()

    def shape(self, layer):
        return (9, self.WIDTHS[layer + 1], self.WIDTHS[layer], 4, 4)

    def init_neumann(self, kappa, noise=0.0, seed=1):
        """sum_{k<=4} (kappa H)^k, the truncated Neumann series of D^-1."""
        self.ctx.check(self._L().ii_net_init(self.ctx.handle, float(kappa), float(noise),
                                             int(seed)), "ii_net_init")

    def set_weights(self, layer, w):
        w = np.ascontiguousarray(np.asarray(w, dtype=np.complex64).reshape(self.shape(layer)))
        f = w.view(np.float32).ravel()
        self.ctx.check(self._L().ii_net_set_weights(self.ctx.handle, int(layer), _lib.fptr(f)),
                       "ii_net_set_weights")

    def get_weights(self, layer):
        shp = self.shape(layer)
        f = np.empty(2 * int(np.prod(shp)), np.float32)
        self.ctx.check(self._L().ii_net_get_weights(self.ctx.handle, int(layer), _lib.fptr(f)),
                       "ii_net_get_weights")
        return f.view(np.complex64).reshape(shp)

    def layer_apply(self, layer, x, adjoint=False):
        cin, cout = self.WIDTHS[layer], self.WIDTHS[layer + 1]
        if adjoint:
            cin, cout = cout, cin
        x = _lib.as_f64(x, cin * 24 * self.ctx.V)
        out = np.empty(cout * 24 * self.ctx.V)
        self.ctx.check(self._L().ii_net_layer_apply(self.ctx.handle, int(layer), int(adjoint),
                                                    _lib.dptr(x), _lib.dptr(out)),
                       "ii_net_layer_apply")
        return out

    def layer_wgrad(self, layer, x, g):
        cin, cout = self.WIDTHS[layer], self.WIDTHS[layer + 1]
        x = _lib.as_f64(x, cin * 24 * self.ctx.V)
        g = _lib.as_f64(g, cout * 24 * self.ctx.V)
        shp = self.shape(layer)
        f = np.empty(2 * int(np.prod(shp)), np.float32)
        self.ctx.check(self._L().ii_net_layer_wgrad(self.ctx.handle, int(layer), _lib.dptr(x),
                                                    _lib.dptr(g), _lib.fptr(f)),
                       "ii_net_layer_wgrad")
        return f.view(np.complex64).reshape(shp)

    def loss_grad(self, bs, ys):
        """Loss sum_k |M b_k - y_k|^2/|y_k|^2 and its gradient per layer."""
        nh = 24 * self.ctx.V // 2
        b = _lib.as_f64(np.concatenate([np.ravel(v) for v in bs]), nh * len(bs))
        y = _lib.as_f64(np.concatenate([np.ravel(v) for v in ys]), nh * len(ys))
        sizes = [int(np.prod(self.shape(l))) for l in range(_lib.NET_LAYERS)]
        f = np.empty(2 * sum(sizes), np.float32)
        loss = ctypes.c_double()
        self.ctx.check(self._L().ii_net_loss_grad(self.ctx.handle, len(bs), _lib.dptr(b),
                                                  _lib.dptr(y), ctypes.byref(loss), _lib.fptr(f)),
                       "ii_net_loss_grad")
        c = f.view(np.complex64)
        grads, off = [], 0
        for l, n in enumerate(sizes):
            grads.append(c[off:off + n].reshape(self.shape(l)))
            off += n
        return loss.value, grads

    def apply(self, r):
        r = _lib.as_f64(r, 24 * self.ctx.V // 2)
        z = np.empty_like(r)
        self.ctx.check(self._L().ii_net_apply(self.ctx.handle, _lib.dptr(r), _lib.dptr(z)),
                       "ii_net_apply")
        return z

    def bench_batched(self, nrhs, nrep=5):
        """Average ms of the four stage kernels of apply_batched."""
        ms = np.zeros(4)
        self.ctx.check(self._L().ii_bench_net_batched(self.ctx.handle, int(nrhs), int(nrep),
                                                      _lib.dptr(ms)), "ii_bench_net_batched")
        return ms

    def tune(self, nrhs, force=None):
        """Chooses (or, with force = 4 group sizes, sets) how many rhs each
        stage kernel of apply_batched processes side by side per tile;
        returns (group sizes, report)."""
        f = None
        if force is not None:
            f = (ctypes.c_int * 4)(*[int(v) for v in force])
        rg = (ctypes.c_int * 4)()
        buf = ctypes.create_string_buffer(1024)
        self.ctx.check(self._L().ii_net_tune(self.ctx.handle, int(nrhs), f, rg, buf, 1024),
                       "ii_net_tune")
        return list(rg), buf.value.decode()

    def apply_batched(self, r):
        """z_k = M r_k for a stack of even half fields, all rhs per launch
        (the evaluation of the batched flexible solver)."""
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1, 24 * self.ctx.V // 2))
        z = np.empty_like(r)
        self.ctx.check(self._L().ii_net_apply_batched(self.ctx.handle, r.shape[0], _lib.dptr(r),
                                                      _lib.dptr(z)), "ii_net_apply_batched")
        return z

    def train(self, kappa, npairs=64, epochs=20, batch=4, lr=1e-3, pair_tol=1e-8, seed=1,
              keep_weights=False):
        p = _lib.TrainParamsC(int(npairs), int(epochs), int(batch), float(lr), float(pair_tol),
                              int(seed), int(keep_weights))
        rep = _lib.TrainReportC()
        self.ctx.check(self._L().ii_net_train(self.ctx.handle, float(kappa), ctypes.byref(p),
                                              ctypes.byref(rep)), "ii_net_train")
        return {"loss_first": rep.loss_first, "loss_last": rep.loss_last,
                "pairs_ms": rep.pairs_ms, "train_ms": rep.train_ms, "steps": rep.steps}

    def bench(self, nrep=10):
        """Average device time (ms) of each forward layer."""
        ms = np.zeros(_lib.NET_LAYERS)
        self.ctx.check(self._L().ii_net_bench(self.ctx.handle, int(nrep), _lib.dptr(ms)),
                       "ii_net_bench")
        return ms

    def save(self, path):
        self.ctx.check(self._L().ii_net_save(self.ctx.handle, str(path).encode()), "ii_net_save")

    def load(self, path):
        self.ctx.check(self._L().ii_net_load(self.ctx.handle, str(path).encode()), "ii_net_load")


class Solver_FGCR:
    """Flexible GCR on Dhat x = b with the learned preconditioner (or none),
    FP64 vectors, restart every nkrylov directions; plain CG on
    Dhat^dag Dhat takes over if the flexible iteration stalls."""

    def __init__(self, ctx, kappa, tol=1e-10, maxiter=1000, nkrylov=16, use_net=True):
        self.ctx, self.kappa, self.tol = ctx, float(kappa), float(tol)
        self.maxiter, self.nkrylov, self.use_net = int(maxiter), int(nkrylov), bool(use_net)

    def solve(self, x, b):
        b = _lib.as_f64(b, 24 * self.ctx.V // 2)
        out = _lib.SolverOutputC()
        xx = np.empty_like(b)
        st = _lib.load().ii_fgcr_solve(self.ctx.handle, self.kappa, _lib.dptr(xx), _lib.dptr(b),
                                       self.tol, self.maxiter, self.nkrylov, int(self.use_net),
                                       ctypes.byref(out))
        self.ctx.check(st, "ii_fgcr_solve")
        x[:] = xx
        return SolverOutput(out)
