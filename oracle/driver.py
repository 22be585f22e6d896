"""Step driver for the CPU checkers -- TEST INFRASTRUCTURE ONLY.

Runs an ``FDTDPlan`` through ``fdtd_oracle.c`` (float64 oracle, or the float32
OpenMP build used as the CPU baseline), in the reference's per-step order
(witwin/maxwell/fdtd/runtime/stepping.py:1740-1815, :2289-2334):

    H update (+CPML psi)             magnetic.cu:62-117
    H surface terms at t             stepping.py:1776-1777
    ADE advance from E^n             stepping.py:1794
    E update (+CPML psi)             electric.cu:463-553
    E surface terms at t + dt/2      stepping.py:2300-2307
    additive point sources at t      stepping.py:2309-2310
    ADE subtraction                  stepping.py:2316 (materials.py:783-798)
    full-field DFT                   spectral.py:627-652
    observers                        observers.py:577-659
    TFSF auxiliary line (numpy)      sources.cu:803-839

Pins (beyond those of fdtd_oracle.c): the auxiliary line by
tests/test_tfsf_oracle.py (the 3-D run of a transversely uniform field equals
the line in the total-field region and vanishes in the scattered-field region);
monitor payloads and fluxes by the pins of oracle/monitors.py; the sparse pole
update (state touched only where a pole acts) by
tests/test_oracle_pinning.py::test_sparse_pole_update_equals_dense.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
_LIBS = {}

_COMP = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
_KIND = {"debye": 0, "drude": 1, "lorentz": 2}


def build(force=False):
    """Compile the checkers with the committed Makefile."""
    targets = [os.path.join(BUILD, "liboracle_f64.so"), os.path.join(BUILD, "libcpu_f32.so")]
    if force or not all(os.path.exists(t) for t in targets):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return targets


def load(precision="f64"):
    if precision in _LIBS:
        return _LIBS[precision]
    name = "liboracle_f64.so" if precision == "f64" else "libcpu_f32.so"
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    real = ctypes.c_double if precision == "f64" else ctypes.c_float
    P = ctypes.POINTER(real)
    PP = ctypes.POINTER(P)
    lib.or_update_h.argtypes = [ctypes.c_void_p, PP, PP, PP]
    lib.or_update_e.argtypes = [ctypes.c_void_p, PP, PP, PP]
    i64, i32, dbl = ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    lib.or_ade_advance.argtypes = [ctypes.c_int, i64, P, P, dbl, dbl, dbl, P, P]
    lib.or_ade_apply.argtypes = [i64, P, P, dbl, P]
    IP = ctypes.POINTER(i64)
    lib.or_ade_advance_idx.argtypes = [ctypes.c_int, i64, IP, P, P, dbl, dbl, dbl, P, P]
    lib.or_ade_apply_idx.argtypes = [i64, IP, P, P, dbl, P]
    lib.or_add_patch.argtypes = [P, i32, i32, P, i32, i32, i32, i32, i32, i32, dbl]
    lib.or_add_shifted.argtypes = [P, i32, i32, P, ctypes.POINTER(dbl), i32, i32, i32, i32, i32, i32, dbl, i32,
                                   dbl, dbl, dbl, dbl, dbl]
    lib.or_accumulate.argtypes = [i64, P, i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl), P, P]
    lib.or_reverse_step.argtypes = [ctypes.c_void_p, PP, PP, PP, PP, ctypes.c_void_p, PP, PP, PP, PP, PP]
    lib.or_max_abs.argtypes = [i64, P]
    lib.or_max_abs.restype = dbl
    lib._real = real
    lib._dtype = np.float64 if precision == "f64" else np.float32
    _LIBS[precision] = lib
    return lib


class _Grid(ctypes.Structure):
    pass


def _grid_struct(real):
    P = ctypes.POINTER(real)
    fields = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
              ("bc_lo", ctypes.c_int32 * 3), ("bc_hi", ctypes.c_int32 * 3), ("cpml", ctypes.c_int32),
              ("pad", ctypes.c_int32), ("inv_e", P * 3), ("inv_h", P * 3), ("decay", P * 6), ("curl", P * 6),
              ("decay_value", real * 6), ("curl_value", real * 6), ("kinv_e", P * 3), ("b_e", P * 3), ("c_e", P * 3), ("kinv_h", P * 3), ("b_h", P * 3), ("c_h", P * 3)]
    return type("OrGrid", (ctypes.Structure,), {"_fields_": fields})


def _ade_adj_struct(real):
    P = ctypes.POINTER(real)
    fields = [("count", ctypes.c_int32 * 3), ("kind", (ctypes.c_int32 * 4) * 3), ("decay", (ctypes.c_double * 4) * 3),
              ("restoring", (ctypes.c_double * 4) * 3), ("drive", (P * 4) * 3), ("apol", (P * 4) * 3),
              ("acur", (P * 4) * 3), ("inv_eps", P * 3), ("dt", ctypes.c_double)]
    return type("OrAdeAdj", (ctypes.Structure,), {"_fields_": fields})


def _uniform(a):
    """The value of an array that holds one value everywhere, else None."""
    a = np.asarray(a)
    if a.size == 0:
        return None
    v = a.flat[0]
    if all(st == 0 for st in a.strides) or bool(np.all(a == v)):
        return float(v)
    return None


def _iptr(arr):
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _ptr(arr, real):
    return arr.ctypes.data_as(ctypes.POINTER(real))


class OracleRun:
    """Holds the state of one plan being stepped on the CPU."""

    def __init__(self, plan, precision="f64", record_monitors=None):
        self.plan = plan
        self.lib = load(precision)
        dt_np = self.lib._dtype
        real = self.lib._real
        self._keep = []

        def arr(a):
            out = np.ascontiguousarray(np.asarray(a, dtype=dt_np))
            self._keep.append(out)
            return out

        nx, ny, nz = plan.shape
        self.fields = {c: np.zeros(plan.field_shape(c), dtype=dt_np) for c in _COMP}
        self.psi = [np.zeros(plan.field_shape(c), dtype=dt_np) for c in
                    ("Ex", "Ex", "Ey", "Ey", "Ez", "Ez", "Hx", "Hx", "Hy", "Hy", "Hz", "Hz")]
        G = _grid_struct(real)
        g = G()
        g.nx, g.ny, g.nz = nx, ny, nz
        for a in range(3):
            g.bc_lo[a] = plan.bc_lo[a]
            g.bc_hi[a] = plan.bc_hi[a]
            g.inv_e[a] = _ptr(arr(plan.inv_e[a]), real)
            g.inv_h[a] = _ptr(arr(plan.inv_h[a]), real)
        for ci, c in enumerate(_COMP):
            dec, curl = plan.coef[c]
            for a, ptrs, values in ((dec, g.decay, g.decay_value), (curl, g.curl, g.curl_value)):
                u = _uniform(a)
                if u is None:
                    ptrs[ci] = _ptr(arr(np.broadcast_to(a, plan.field_shape(c))), real)
                else:  # one value over the whole component: passed as a scalar
                    ptrs[ci] = None
                    values[ci] = u
        g.cpml = 1 if plan.uses_cpml else 0
        if plan.uses_cpml:
            for a, ax in enumerate("xyz"):
                p = plan.cpml[ax]
                g.kinv_e[a] = _ptr(arr(1.0 / p["kappa_e"]), real)
                g.b_e[a] = _ptr(arr(p["b_e"]), real)
                g.c_e[a] = _ptr(arr(p["c_e"]), real)
                g.kinv_h[a] = _ptr(arr(1.0 / p["kappa_h"]), real)
                g.b_h[a] = _ptr(arr(p["b_h"]), real)
                g.c_h[a] = _ptr(arr(p["c_h"]), real)
        self.grid = g
        P = ctypes.POINTER(real)
        self._eptr = (P * 3)(*[_ptr(self.fields[c], real) for c in ("Ex", "Ey", "Ez")])
        self._hptr = (P * 3)(*[_ptr(self.fields[c], real) for c in ("Hx", "Hy", "Hz")])
        self._psiptr = (P * 12)(*[_ptr(p, real) for p in self.psi])

        self.ade = {}
        for c in ("Ex", "Ey", "Ez"):
            entries = []
            for pole in plan.ade[c]["poles"]:
                drive = arr(pole["drive"])
                # the samples where the pole acts (its state is zero elsewhere)
                idx = np.ascontiguousarray(np.flatnonzero(drive), dtype=np.int64)
                entries.append({"kind": _KIND[pole["kind"]], "decay": pole["decay"], "restoring": pole["restoring"],
                                "drive": drive, "idx": idx, "pol": np.zeros(plan.field_shape(c), dt_np),
                                "cur": np.zeros(plan.field_shape(c), dt_np)})
            self.ade[c] = {"inv_eps": arr(plan.ade[c]["inv_eps"]), "poles": entries}

        self.point_terms = [(t["field"], t["offset"], arr(t["patch"])) for t in plan.point_terms]
        self.e_terms = [(t, arr(t["patch"]), np.ascontiguousarray(t["delay"], dtype=np.float64)) for t in plan.e_terms]
        self.h_terms = [(t, arr(t["patch"]), np.ascontiguousarray(t["delay"], dtype=np.float64)) for t in plan.h_terms]

        # TFSF auxiliary 1-D line (float64; advanced by _line_h / _line_e below)
        self.aux = plan.aux
        if self.aux is not None:
            self.aux_e = np.zeros(self.aux["n_e"])
            self.aux_h = np.zeros(self.aux["n_e"] - 1)

        self.obs = plan.observers
        # the oracle's own monitor sampling and payload assembly (oracle/monitors.py)
        self.monitors = None
        if record_monitors is None:
            record_monitors = precision == "f64"
        if self.obs is not None and record_monitors:
            from .monitors import OracleObservers
            self.monitors = OracleObservers(plan)
        if self.obs is not None:
            nf = len(self.obs.frequencies)
            self.point_re = np.zeros((nf, len(self.obs.points)))
            self.point_im = np.zeros((nf, len(self.obs.points)))
            shapes = {c: plan.field_shape(c) for c in _COMP}
            self.plane_re = [np.zeros((nf,) + self.obs.plane_shape(shapes, k)) for k in self.obs.planes]
            self.plane_im = [np.zeros((nf,) + self.obs.plane_shape(shapes, k)) for k in self.obs.planes]
        self.dft = None
        if plan.dft_frequencies:
            nf = len(plan.dft_frequencies)
            self.dft = {c: (np.zeros((nf,) + plan.field_shape(c), dt_np), np.zeros((nf,) + plan.field_shape(c), dt_np))
                        for c in ("Ex", "Ey", "Ez")}
        self.step_index = 0
        # running peak |field| per component over all steps taken, the scale
        # against which late-time (decayed) fields are judged
        self.peak = {c: 0.0 for c in _COMP}
        self.track_peak = precision == "f64"
        self.trajectory = None  # list of per-step forward states when recording for the adjoint

    # ------------------------------------------------------------------ steps
    def _shifted(self, terms, time):
        real = self.lib._real
        for t, patch, delay in terms:
            f = self.fields[t["field"]]
            st = t["st"]
            self.lib.or_add_shifted(_ptr(f, real), f.shape[1], f.shape[2], _ptr(patch, real),
                                    delay.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), *patch.shape,
                                    *t["offset"], time, st["kind_code"], st["frequency"], st["fwidth"],
                                    st["amplitude"], st["phase"], st["delay"])

    def _line_terms(self, terms, line):
        """field[patch] += patch * line[base + step * p_axis] (TFSF faces)."""
        for t in terms:
            f = self.fields[t["field"]]
            patch = t["patch"]
            o = t["offset"]
            box = tuple(slice(o[a], o[a] + patch.shape[a]) for a in range(3))
            if t["line_axis"] < 0:
                samples = line[t["line_base"]]
            else:
                ax = t["line_axis"]
                idx = t["line_base"] + t["line_step"] * np.arange(patch.shape[ax])
                shape = [1, 1, 1]
                shape[ax] = patch.shape[ax]
                samples = line[idx].reshape(shape)
            f[box] += (patch * samples).astype(f.dtype)

    def step(self):
        plan, lib, real = self.plan, self.lib, self.lib._real
        n = self.step_index
        t = n * plan.dt
        e_before = {c: self.fields[c].copy() for c in _COMP[:3]} if self.trajectory is not None else None
        lib.or_update_h(ctypes.byref(self.grid), self._eptr, self._hptr, self._psiptr)
        self._shifted(self.h_terms, t)
        if self.aux is not None:
            self._line_terms(plan.line_h_terms, self.aux_e)
            self._line_h()
        for c in ("Ex", "Ey", "Ez"):
            e = self.fields[c]
            for p in self.ade[c]["poles"]:
                lib.or_ade_advance_idx(p["kind"], p["idx"].size, _iptr(p["idx"]), _ptr(e, real),
                                       _ptr(p["drive"], real), p["decay"], p["restoring"], plan.dt,
                                       _ptr(p["pol"], real), _ptr(p["cur"], real))
        lib.or_update_e(ctypes.byref(self.grid), self._eptr, self._hptr, self._psiptr)
        if self.aux is not None:
            self._line_terms(plan.line_e_terms, self.aux_h)
            self._line_e(n)
        self._shifted(self.e_terms, t + 0.5 * plan.dt)
        for ti, (comp, off, patch) in enumerate(self.point_terms):
            f = self.fields[comp]
            lib.or_add_patch(_ptr(f, real), f.shape[1], f.shape[2], _ptr(patch, real), *patch.shape, *off,
                             float(plan.signals[ti, n]))
        for c in ("Ex", "Ey", "Ez"):
            e = self.fields[c]
            for p in self.ade[c]["poles"]:
                lib.or_ade_apply_idx(p["idx"].size, _iptr(p["idx"]), _ptr(p["cur"], real),
                                     _ptr(self.ade[c]["inv_eps"], real), plan.dt, _ptr(e, real))
        if self.dft is not None:
            wc = np.ascontiguousarray(plan.dft_weights[0][n])
            ws = np.ascontiguousarray(plan.dft_weights[1][n])
            for c, (re, im) in self.dft.items():
                lib.or_accumulate(self.fields[c].size, _ptr(self.fields[c], real), len(wc),
                                  wc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  ws.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(re, real), _ptr(im, real))
        if self.obs is not None:
            self._observe(n)
            if self.monitors is not None:
                self.monitors.record(n, self.fields)
        if self.track_peak:
            for c in _COMP:
                f = self.fields[c]
                self.peak[c] = max(self.peak[c], lib.or_max_abs(f.size, _ptr(f, real)))
        if self.trajectory is not None:
            # E^n, H^{n+1/2} and the E-side psi^{n+1}: what the coefficient
            # gradients of step n need on the way back
            self.trajectory.append((e_before, {c: self.fields[c].copy() for c in _COMP[3:]},
                                    [p.copy() for p in self.psi[:6]]))
        self.step_index += 1

    def _line_h(self):
        """Auxiliary line H: h[m] <- decay_h[m] h[m] - curl_h[m] (e[m+1] - e[m])
        (kernels/sources.cu:803-814)."""
        a, e, h = self.aux, self.aux_e, self.aux_h
        h *= a["decay_h"]
        h -= a["curl_h"] * np.diff(e)

    def _line_e(self, n):
        """Auxiliary line E: hard source at sample 0, zero at the far end, the 1-D
        curl update in between (kernels/sources.cu:816-839)."""
        a, e, h = self.aux, self.aux_e, self.aux_h
        inner = slice(1, e.size - 1)
        e[inner] = a["decay_e"][inner] * e[inner] - a["curl_e"][inner] * np.diff(h)
        e[0] = a["source"][n]
        e[-1] = 0.0

    def _observe(self, n):
        ec, es, hc, hs = (w[n] for w in self.plan.obs_weights)
        for pi, (comp, idx) in enumerate(self.obs.points):
            v = float(self.fields[comp][idx])
            wc, ws = (ec, es) if comp[0] == "E" else (hc, hs)
            self.point_re[:, pi] += v * wc
            self.point_im[:, pi] += v * ws
        for qi, (comp, axis, index) in enumerate(self.obs.planes):
            f = self.fields[comp]
            sl = np.take(f, index, axis="xyz".index(axis)).astype(np.float64)
            wc, ws = (ec, es) if comp[0] == "E" else (hc, hs)
            self.plane_re[qi] += wc[:, None, None] * sl[None]
            self.plane_im[qi] += ws[:, None, None] * sl[None]

    def run(self, steps=None):
        steps = self.plan.time_steps if steps is None else steps
        for _ in range(steps):
            self.step()
        return self

    def monitor_results(self):
        """Monitor payloads assembled by the oracle's own code (oracle/monitors.py):
        per monitor, the component data (point values or plane arrays with their
        coordinates) and, for flux planes, the flux."""
        return self.monitors.results() if self.monitors is not None else {}


def run_plan(plan, steps=None, precision="f64"):
    return OracleRun(plan, precision).run(steps)


class OracleAdjoint:
    """Reverse sweep of a recorded ``OracleRun`` (float64): the exact transpose of
    the forward steps (fdtd_oracle.c ``or_reverse_step``), seeded by the
    gradients of the observer sums (the transpose of ``_observe`` and the
    full-field DFT), accumulating dL/d(decay_E) and dL/d(curl_E).

    Follows the reverse phases of fdtd/adjoint/native.py:48-727 and the seed
    injection of fdtd/adjoint/seeds.py:922-990."""

    def __init__(self, run: OracleRun):
        self.run = run
        plan, lib = run.plan, run.lib
        real = lib._real
        dt_np = lib._dtype
        self.ae = [np.zeros(plan.field_shape(c), dt_np) for c in _COMP[:3]]
        self.ah = [np.zeros(plan.field_shape(c), dt_np) for c in _COMP[3:]]
        self.apsi = [np.zeros_like(p) for p in run.psi]
        self.tmp = [np.zeros_like(a) for a in self.ae]
        self.g_decay = [np.zeros_like(a) for a in self.ae]
        self.g_curl = [np.zeros_like(a) for a in self.ae]
        P = ctypes.POINTER(real)
        self._pp = lambda arrs: (P * len(arrs))(*[_ptr(a, real) for a in arrs])
        self.adj_poles = {}
        A = _ade_adj_struct(real)
        ad = A()
        ad.dt = plan.dt
        any_pole = False
        for ci, c in enumerate(_COMP[:3]):
            entry = run.ade[c]
            ad.inv_eps[ci] = _ptr(entry["inv_eps"], real)
            poles = []
            for pi, p in enumerate(entry["poles"]):
                any_pole = True
                apol = np.zeros(plan.field_shape(c), dt_np)
                acur = np.zeros(plan.field_shape(c), dt_np)
                poles.append({"pol": apol, "cur": acur})
                ad.kind[ci][pi] = p["kind"]
                ad.decay[ci][pi] = p["decay"]
                ad.restoring[ci][pi] = p["restoring"]
                ad.drive[ci][pi] = _ptr(p["drive"], real)
                ad.apol[ci][pi] = _ptr(apol, real)
                ad.acur[ci][pi] = _ptr(acur, real)
            ad.count[ci] = len(poles)
            self.adj_poles[c] = poles
        self.ad = ad if any_pole else None

    # transpose of OracleRun._observe and the full-field DFT at step n
    def seed(self, n, grads):
        plan, run = self.run.plan, self.run
        fields = dict(zip(_COMP, self.ae + self.ah))
        if run.obs is not None and plan.obs_weights is not None:
            ec, es, hc, hs = (w[n] for w in plan.obs_weights)
            g_pre, g_pim = grads.get("point_re"), grads.get("point_im")
            for pi, (comp, idx) in enumerate(run.obs.points):
                wc, ws = (ec, es) if comp[0] == "E" else (hc, hs)
                v = 0.0
                if g_pre is not None:
                    v += float(np.dot(g_pre[:, pi], wc))
                if g_pim is not None:
                    v += float(np.dot(g_pim[:, pi], ws))
                fields[comp][idx] += v
            for qi, (comp, axis, index) in enumerate(run.obs.planes):
                wc, ws = (ec, es) if comp[0] == "E" else (hc, hs)
                acc = 0.0
                if grads.get("plane_re") is not None:
                    acc = acc + np.tensordot(wc, grads["plane_re"][qi], axes=(0, 0))
                if grads.get("plane_im") is not None:
                    acc = acc + np.tensordot(ws, grads["plane_im"][qi], axes=(0, 0))
                sl = [slice(None)] * 3
                sl["xyz".index(axis)] = index
                fields[comp][tuple(sl)] += acc
        if plan.dft_weights is not None and grads.get("dft") is not None:
            wc, ws = plan.dft_weights[0][n], plan.dft_weights[1][n]
            for ci, c in enumerate(_COMP[:3]):
                gre, gim = grads["dft"][c]
                self.ae[ci] += np.tensordot(wc, gre, axes=(0, 0)) + np.tensordot(ws, gim, axes=(0, 0))

    def reverse_step(self, n, gradients=True):
        run = self.run
        real = run.lib._real
        fe = fh = fpsi = None
        if gradients:
            e_n, h_half, psi_e = run.trajectory[n]
            fe = self._pp([e_n[c] for c in _COMP[:3]])
            fh = self._pp([h_half[c] for c in _COMP[3:]])
            fpsi = self._pp(psi_e)
        run.lib.or_reverse_step(ctypes.byref(run.grid), self._pp(self.ae), self._pp(self.ah), self._pp(self.apsi),
                                self._pp(self.tmp), ctypes.byref(self.ad) if self.ad is not None else None,
                                fe, fh, fpsi, self._pp(self.g_decay) if gradients else None,
                                self._pp(self.g_curl) if gradients else None)

    def run_reverse(self, grads):
        steps = len(self.run.trajectory)
        for n in range(steps - 1, -1, -1):
            self.seed(n, grads)
            self.reverse_step(n)
        return self
