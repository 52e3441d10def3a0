"""ctypes wrapper of oracle/liboracle.so, the CPU restatement of the
IroIro++ Wilson even/odd solve path.

TEST INFRASTRUCTURE ONLY.  Importable from tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py, where it is the checker.  The product
(iroiro_amd) never imports it.

Arrays are float64 numpy arrays in the reference layouts (Format_F spinors,
Format_G links, even/odd half fields in esec/osec order); see
oracle/wilson_oracle.h for the citations.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        c_int, c_long, c_double, c_u64 = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_uint64
        vp = ctypes.c_void_p
        sig = {
            "or_geom_new": (vp, [_ip]),
            "or_geom_free": (None, [vp]),
            "or_geom_volume": (c_long, [vp]),
            "or_geom_eo_tables": (None, [vp, _ip, _ip]),
            "or_apbc": (None, [vp, _dp]),
            "or_gamma5": (None, [c_long, _dp, _dp]),
            "or_hop_eo": (None, [vp, _dp, _dp, _dp, c_int, c_int, c_double]),
            "or_dirac_eo": (None, [vp, _dp, _dp, _dp, c_int, c_double]),
            "or_ddagd": (None, [vp, _dp, _dp, _dp, c_double]),
            "or_dirac_full": (None, [vp, _dp, _dp, _dp, c_int, c_double]),
            "or_md_force_full": (None, [vp, _dp, _dp, _dp, _dp, c_double]),
            "or_md_force_pm_full": (None, [vp, _dp, _dp, _dp, _dp, c_int]),
            "or_md_force_eo": (None, [vp, _dp, _dp, _dp, _dp, c_double]),
            "or_traceless_antihermite": (None, [c_long, _dp, _dp]),
            "or_cg": (c_int, [vp, _dp, c_double, _dp, _dp, c_double, c_int,
                              ctypes.POINTER(c_int), ctypes.POINTER(c_double)]),
            "or_multishift_cg": (c_int, [vp, _dp, c_double, _dp, c_int, _dp, _dp, c_double,
                                         c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_double),
                                         ctypes.POINTER(c_double)]),
            "or_eo_invert": (c_int, [vp, _dp, c_double, _dp, _dp, c_double, c_int,
                                     ctypes.POINTER(c_int), ctypes.POINTER(c_double)]),
            "or_true_residual_full": (c_double, [vp, _dp, c_double, _dp, _dp]),
            "or_reunit": (None, [_dp]),
            "or_gauge_expH": (None, [vp, c_double, c_u64, _dp]),
            "or_gauss_field": (None, [c_long, c_u64, _dp]),
            "or_plaquette": (c_double, [vp, _dp]),
            "or_heatbath_sweep": (None, [vp, _dp, c_double, c_u64, c_int]),
            "or_uniform": (c_double, [c_u64, c_u64, c_u64]),
            "or_net_layer": (None, [vp, _dp, c_int, c_int, _dp, _dp, _dp, c_int]),
            "or_net_layer_adj": (None, [vp, _dp, c_int, c_int, _dp, _dp, _dp, c_int]),
            "or_net_layer_wgrad": (None, [vp, _dp, c_int, c_int, _dp, _dp, _dp, c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def kappa_from_mass(mass):
    """Dirac_Wilson: kpp_ = 0.5/(4.0+mass) (lib/Dirac_ops/dirac_wilson.hpp:93)."""
    return 0.5 / (4.0 + mass)


class Lattice:
    """Geometry of one lattice (lexicographic and even/odd tables)."""

    def __init__(self, dims):
        self.dims = tuple(int(d) for d in dims)
        if self.dims[0] % 2:
            raise ValueError("Lx must be even (siteIndex_EvenOdd.hpp:45)")
        self._g = lib().or_geom_new(np.asarray(self.dims, dtype=np.int32))
        self.V = int(np.prod(self.dims))
        self.Vh = self.V // 2

    def __del__(self):
        g = getattr(self, "_g", None)
        if g:
            lib().or_geom_free(g)
            self._g = None

    # sizes in doubles
    @property
    def gauge_size(self):
        return 18 * 4 * self.V

    @property
    def full_size(self):
        return 24 * self.V

    @property
    def half_size(self):
        return 24 * self.Vh

    def eo_tables(self):
        e = np.empty(self.Vh, np.int32)
        o = np.empty(self.Vh, np.int32)
        lib().or_geom_eo_tables(self._g, e, o)
        return e, o

    def split(self, full):
        e, o = self.eo_tables()
        f = full.reshape(self.V, 24)
        return np.ascontiguousarray(f[e].ravel()), np.ascontiguousarray(f[o].ravel())

    def merge(self, fe, fo):
        e, o = self.eo_tables()
        out = np.empty((self.V, 24))
        out[e] = fe.reshape(self.Vh, 24)
        out[o] = fo.reshape(self.Vh, 24)
        return out.ravel()

    # ---- operators
    def apbc(self, u):
        u = np.array(u, dtype=np.float64, copy=True)
        lib().or_apbc(self._g, u)
        return u

    def hop_eo(self, u, f, eotype, dag, kappa):
        w = np.empty(self.half_size)
        lib().or_hop_eo(self._g, u, w, np.ascontiguousarray(f), int(eotype), int(dag), kappa)
        return w

    def dirac_eo(self, u, f, dag, kappa):
        w = np.empty(self.half_size)
        lib().or_dirac_eo(self._g, u, w, np.ascontiguousarray(f), int(dag), kappa)
        return w

    def ddagd(self, u, f, kappa):
        w = np.empty(self.half_size)
        lib().or_ddagd(self._g, u, w, np.ascontiguousarray(f), kappa)
        return w

    def dirac_full(self, u, f, dag, kappa):
        w = np.empty(self.full_size)
        lib().or_dirac_full(self._g, u, w, np.ascontiguousarray(f), int(dag), kappa)
        return w

    def md_force_full(self, u, eta, zeta, kappa):
        f = np.empty(self.gauge_size)
        lib().or_md_force_full(self._g, u, f, np.ascontiguousarray(eta),
                               np.ascontiguousarray(zeta), kappa)
        return f

    def md_force_pm_full(self, u, fce, eta, zeta, minus):
        """Accumulates md_force_p (minus=0) or md_force_m (minus=1) into fce."""
        lib().or_md_force_pm_full(self._g, u, fce, np.ascontiguousarray(eta),
                                  np.ascontiguousarray(zeta), int(minus))
        return fce

    def md_force_eo(self, u, eta, zeta, kappa):
        f = np.empty(self.gauge_size)
        lib().or_md_force_eo(self._g, u, f, np.ascontiguousarray(eta),
                             np.ascontiguousarray(zeta), kappa)
        return f

    def traceless_antihermite(self, m):
        out = np.empty(self.gauge_size)
        lib().or_traceless_antihermite(self.gauge_size // 18, out, np.ascontiguousarray(m))
        return out

    def cg(self, u, kappa, b, prec, maxiter):
        x = np.empty(self.half_size)
        it = ctypes.c_int()
        diff = ctypes.c_double()
        rc = lib().or_cg(self._g, u, kappa, x, np.ascontiguousarray(b), prec, maxiter,
                         ctypes.byref(it), ctypes.byref(diff))
        return x, {"converged": rc == 0, "iterations": it.value, "diff": diff.value}

    def multishift_cg(self, u, kappa, shifts, b, prec, maxiter):
        shifts = np.ascontiguousarray(shifts, dtype=np.float64)
        x = np.empty(len(shifts) * self.half_size)
        nconv = ctypes.c_int()
        diff = ctypes.c_double()
        odiff = ctypes.c_double()
        rc = lib().or_multishift_cg(self._g, u, kappa, shifts, len(shifts), x,
                                    np.ascontiguousarray(b), prec, maxiter, ctypes.byref(nconv),
                                    ctypes.byref(diff), ctypes.byref(odiff))
        return x.reshape(len(shifts), self.half_size), {
            "converged": rc == 0, "nconv": nconv.value, "diff": diff.value,
            "out_diff": odiff.value}

    def eo_invert(self, u, kappa, src, prec, maxiter):
        sol = np.empty(self.full_size)
        it = ctypes.c_int()
        diff = ctypes.c_double()
        rc = lib().or_eo_invert(self._g, u, kappa, sol, np.ascontiguousarray(src), prec,
                                maxiter, ctypes.byref(it), ctypes.byref(diff))
        return sol, {"converged": rc == 0, "iterations": it.value, "diff": diff.value}

    def true_residual_full(self, u, kappa, x, b):
        return lib().or_true_residual_full(self._g, u, kappa, np.ascontiguousarray(x),
                                           np.ascontiguousarray(b))

    # ---- gauge fields and sources
    def gauge_expH(self, eps, seed):
        u = np.empty(self.gauge_size)
        lib().or_gauge_expH(self._g, eps, seed, u)
        return u

    def plaquette(self, u):
        return lib().or_plaquette(self._g, u)

    def heatbath_sweep(self, u, beta, seed, sweep):
        lib().or_heatbath_sweep(self._g, u, beta, seed, sweep)

    # ---- learned gauge-covariant layers (weights: complex [9][cout][cin][4][4])
    def net_layer(self, u, w, x, round_bf16=False):
        cout, cin = w.shape[1], w.shape[2]
        out = np.empty(cout * self.full_size)
        lib().or_net_layer(self._g, u, cin, cout, _wflat(w), np.ascontiguousarray(x, np.float64),
                           out, int(round_bf16))
        return out

    def net_layer_adj(self, u, w, g, round_bf16=False):
        cout, cin = w.shape[1], w.shape[2]
        out = np.empty(cin * self.full_size)
        lib().or_net_layer_adj(self._g, u, cin, cout, _wflat(w),
                               np.ascontiguousarray(g, np.float64), out, int(round_bf16))
        return out

    def net_layer_wgrad(self, u, cin, cout, x, g, round_bf16=False):
        gw = np.zeros(9 * cout * cin * 32)
        lib().or_net_layer_wgrad(self._g, u, cin, cout, np.ascontiguousarray(x, np.float64),
                                 np.ascontiguousarray(g, np.float64), gw, int(round_bf16))
        return gw.view(np.complex128).reshape(9, cout, cin, 4, 4)

    def unit_gauge(self):
        u = np.zeros((4 * self.V, 18))
        u[:, 0] = u[:, 8] = u[:, 16] = 1.0
        return u.ravel()


def _wflat(w):
    return np.ascontiguousarray(np.asarray(w, dtype=np.complex128)).view(np.float64).ravel()


def bf16_round(a):
    """Round float64 values to the nearest bfloat16 (ties to even)."""
    f = np.asarray(a, dtype=np.float32).copy()
    b = f.view(np.uint32)
    b += np.uint32(0x7FFF) + ((b >> 16) & np.uint32(1))
    b &= np.uint32(0xFFFF0000)
    return b.view(np.float32).astype(np.float64)


def e4m3_round(a):
    """Round to the nearest OCP e4m3 (fp8) value, ties to even, for
    magnitudes below 448: 3 mantissa bits for |a| >= 2^-6, a fixed spacing of
    2^-9 below (the subnormal range)."""
    a = np.asarray(a, dtype=np.float64)
    m = np.abs(a)
    e = np.floor(np.log2(np.where(m > 0, m, 1.0)))
    q = np.where(m >= 2.0 ** -6, 2.0 ** (e - 3), 2.0 ** -9)
    return np.rint(a / q) * q


def fp8_record_round(x, nch, sites):
    """The e4m3 storage of activation records: x holds nch channels of a full
    field ([channel][site][24] doubles); at every site of `sites`, the 8-channel
    groups (channels 8h .. 8h + 7) are scaled by 2^(8 - e), e the exponent of
    their largest magnitude (frexp), rounded to e4m3 and scaled back, so the
    largest magnitude of a group lies in [128, 256) before rounding.  A group
    whose largest magnitude is below 2^-111 keeps scale 1 (its values round
    to zero)."""
    y = np.array(x, dtype=np.float64).reshape(nch, -1, 24)
    for h in range(nch // 8):
        blk = y[8 * h:8 * h + 8][:, sites, :]
        mx = np.abs(blk).max(axis=(0, 2))
        _, e = np.frexp(mx)
        ds = np.where(mx >= 2.0 ** -111, np.ldexp(1.0, e - 8), 1.0)
        y[8 * h:8 * h + 8][:, sites, :] = e4m3_round(blk / ds[None, :, None]) * ds[None, :, None]
    return y.ravel()


class Net:
    """Composition of gauge-covariant layers (linear), widths e.g. (1, 8, 8, 8, 1).
    The preconditioner acts on an even half field b: it is placed on the even
    sites of a full field (odd sites zero), the layers run on the full
    lattice, and the even sites of the output are returned.  Loss of a batch
    of (b_k, y_k) pairs: sum_k |N(b_k) - y_k|^2 / |y_k|^2."""

    def __init__(self, lat, u, weights, round_bf16=False):
        self.lat, self.u, self.w, self.rb = lat, u, list(weights), round_bf16

    def embed(self, b):
        return self.lat.merge(b, np.zeros(self.lat.half_size))

    def forward(self, b):
        acts = [self.embed(b)]
        for w in self.w:
            acts.append(self.lat.net_layer(self.u, w, acts[-1], self.rb))
        return acts

    def apply(self, b):
        return self.lat.split(self.forward(b)[-1])[0]

    def apply_fp8_records(self, b):
        """The network of widths (1, 8, 8, 8, 1) as the device's batched
        evaluation stores it: L1 and L2 on odd sites, L2 on even sites and the
        one-channel mixes Y_d = W4_d L3 + (W4_0 W3_d) L2 (d = 1..8, odd sites;
        z = (W4_0 W3_0) L2 + sum_d T_d Y_d on even sites) rounded to e4m3
        records (fp8_record_round, one scale per site and group of 8 channels
        or 8 paths), everything else exact.  L2(odd) comes from the unrounded
        L1 (the device forms it straight from the input), L2(even) from the
        rounded L1(odd) and the exact L1(even) = W1_0 b."""
        lat, w = self.lat, self.w
        e, o = lat.eo_tables()
        x0 = self.embed(b)
        l1 = lat.net_layer(self.u, w[0], x0, self.rb)
        l2 = lat.net_layer(self.u, w[1], l1, self.rb).reshape(8, lat.V, 24)
        l1q = fp8_record_round(l1, 8, o)
        l2e = lat.net_layer(self.u, w[1], l1q, self.rb).reshape(8, lat.V, 24)
        l2q = np.empty_like(l2)
        l2q[:, e, :] = l2e[:, e, :]
        l2q[:, o, :] = l2[:, o, :]
        l2q = fp8_record_round(fp8_record_round(l2q.ravel(), 8, o), 8, e)
        l3 = lat.net_layer(self.u, w[2], l2q, self.rb)
        w3, w4 = np.asarray(w[2]), np.asarray(w[3])
        zero = np.zeros((9, 8, 8, 4, 4), complex)
        wa, wb = zero.copy(), zero.copy()
        for d in range(1, 9):
            wa[0, d - 1] = w4[d, 0]
            wb[0, d - 1] = np.einsum("kst,kitu->isu", w4[0, 0], w3[d])
        y = (lat.net_layer(self.u, wa, l3, self.rb) + lat.net_layer(self.u, wb, l2q, self.rb))
        yq = fp8_record_round(y, 8, o)
        wi = np.zeros((9, 1, 8, 4, 4), complex)
        wi[0, 0] = np.einsum("kst,kitu->isu", w4[0, 0], w3[0])
        wt = np.zeros((9, 1, 8, 4, 4), complex)
        for d in range(1, 9):
            wt[d, 0, d - 1] = np.eye(4)
        z = lat.net_layer(self.u, wi, l2q, self.rb) + lat.net_layer(self.u, wt, yq, self.rb)
        return lat.split(z)[0]

    def apply_device_form(self, b, rounding=True):
        """The batched device evaluation of the network of widths
        (1, 8, 8, 8, 1) step by step, as ii_bnet.hip computes it, with its
        roundings: the input r, the links and the composite weights rounded
        to bf16 (real and imaginary parts); every transported vector of the
        first three stages rounded to bf16 (the matrix-core operand); L3
        rounded to bf16 before the Y mixes; the records [L1 | L2](odd),
        L2(even) and Y rounded to e4m3 (fp8_record_round).  Sums are exact
        here and FP32 on the device; nothing else differs.
          [L1 | L2](odd) = sum_{d>0} C1_d T_d r,  C1_d = [W1_d ; W2_0 W1_d + W2_d W1_0]
          L2(even) = (W2_0 W1_0) r + sum_{d>0} W2_d T_d L1(odd)
          L3(odd)  = W3_0 L2(odd) + sum_{d>0} W3_d T_d L2(even)
          Y_d      = W4_d L3(odd) + (W4_0 W3_d) L2(odd)
          z(even)  = (W4_0 W3_0) L2(even) + sum_{d>0} T_d Y_d
        With rounding=False every rounding is left out (the same regrouped
        evaluation in exact arithmetic)."""
        lat = self.lat
        e, o = lat.eo_tables()
        b16 = bf16_round if rounding else (lambda a: np.array(a, dtype=np.float64))
        f8 = fp8_record_round if rounding else (lambda x, nch, sites: np.array(x, dtype=np.float64))
        ub = b16(self.u)
        W1, W2, W3, W4 = (np.asarray(w, dtype=np.complex128) for w in self.w)

        def rb(w):
            return b16(w.real) + 1j * b16(w.imag)

        def comp(a, c):  # (a c)[o][i] = sum_k a[o][k] @ c[k][i]
            return np.einsum("okst,kitu->oisu", a, c)

        def layer(w, x, rnd):
            return lat.net_layer(ub, rb(w), x, rnd and rounding)

        x0 = self.embed(b16(b))
        c1 = np.zeros((9, 16, 1, 4, 4), complex)
        for d in range(1, 9):
            c1[d, :8] = W1[d]
            c1[d, 8:] = comp(W2[0], W1[d]) + comp(W2[d], W1[0])
        s1 = f8(layer(c1, x0, True), 16, o).reshape(16, lat.V, 24)
        c2 = np.zeros((9, 8, 8, 4, 4), complex)
        c2[1:] = W2[1:]
        i2 = np.zeros((9, 8, 1, 4, 4), complex)
        i2[0] = comp(W2[0], W1[0])
        l2e = layer(c2, s1[:8].ravel(), True) + layer(i2, x0, False)
        l2q = l2e.reshape(8, lat.V, 24).copy()
        l2q[:, o, :] = s1[8:, o, :]
        l2q = f8(l2q.ravel(), 8, e)
        c3 = np.array(W3)
        c3[0] = 0.0
        i3 = np.zeros_like(W3)
        i3[0] = W3[0]
        l3 = b16(layer(c3, l2q, True) + layer(i3, l2q, False))
        wa = np.zeros((9, 8, 8, 4, 4), complex)
        wb = np.zeros((9, 8, 8, 4, 4), complex)
        for d in range(1, 9):
            wa[0, d - 1] = W4[d, 0]
            wb[0, d - 1] = comp(W4[0], W3[d])[0]
        yq = f8(layer(wa, l3, False) + layer(wb, l2q, False), 8, o)
        wi = np.zeros((9, 1, 8, 4, 4), complex)
        wi[0] = comp(W4[0], W3[0])
        wt = np.zeros((9, 1, 8, 4, 4), complex)
        for d in range(1, 9):
            wt[d, 0, d - 1] = np.eye(4)
        z = layer(wi, l2q, False) + layer(wt, yq, False)
        return lat.split(z)[0]

    def loss_and_grad(self, pairs):
        loss = 0.0
        grads = [np.zeros_like(w) for w in self.w]
        for b, y in pairs:
            acts = self.forward(b)
            e = self.lat.split(acts[-1])[0] - y
            ny = float(np.dot(y, y))
            loss += float(np.dot(e, e)) / ny
            g = self.lat.merge(2.0 * e / ny, np.zeros(self.lat.half_size))
            for li in range(len(self.w) - 1, -1, -1):
                w = self.w[li]
                grads[li] += self.lat.net_layer_wgrad(self.u, w.shape[2], w.shape[1], acts[li], g,
                                                      self.rb)
                if li:
                    g = self.lat.net_layer_adj(self.u, w, g, self.rb)
        return loss, grads


def gamma5(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    w = np.empty_like(f)
    lib().or_gamma5(f.size // 24, w, f)
    return w


def gauss_field(n, seed):
    f = np.empty(n)
    lib().or_gauss_field(n, seed, f)
    return f


def uniform(seed, stream, counter):
    return lib().or_uniform(seed, stream, counter)


def reunit(m):
    m = np.array(m, dtype=np.float64, copy=True)
    lib().or_reunit(m)
    return m
