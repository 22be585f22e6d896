"""``NativeFDTDModule`` with the reference's kernel names, on the C ABI.

The reference reaches its CUDA kernels through
``NativeFDTDModule.<CamelName>(**kwargs).launchRaw()`` (witwin/maxwell/fdtd/cuda/
backend.py:81-115, kernel table :3583-3727).  This module keeps that interface
for the forward hot path -- the same op names, keyword names and argument
meaning -- and binds each op to the per-op entry points of libmaxwell_fdtd.so
(``include/maxwell_fdtd.h``), so reference-side code that drives single ops can
run here unchanged.  Whole steps are better taken through ``mw_run_steps`` (the
single-pass step); these per-op launches are the two-kernel path.

Covered: the H and E updates (standard, dense CPML and compressed-slab CPML),
the Debye/Drude/Lorentz current updates and the polarisation-current subtraction,
additive and time-shifted source patches, the TFSF auxiliary line, the batched
full-field DFT and the point and plane observers.  An op outside that set raises
``AttributeError`` as the reference's does for names it does not implement; an
argument the C ABI has no counterpart for (a time-shifted patch's causal
activation gate) raises ``NotImplementedError`` rather than being ignored.
Every launch goes to torch's current stream.
"""

from __future__ import annotations

import ctypes
from typing import Callable

import torch

from . import native as nat

_AXIS = {"x": 0, "y": 1, "z": 2}
# derivative axes (a, b) of each component's curl, in the order of the psi slots
_CURL_AXES = {"Ex": "yz", "Ey": "xz", "Ez": "xy", "Hx": "yz", "Hy": "xz", "Hz": "xy"}
_COMP = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
_PSI_SLOT = {("Ex", "y"): 0, ("Ex", "z"): 1, ("Ey", "x"): 2, ("Ey", "z"): 3, ("Ez", "x"): 4, ("Ez", "y"): 5,
             ("Hx", "y"): 6, ("Hx", "z"): 7, ("Hy", "x"): 8, ("Hy", "z"): 9, ("Hz", "x"): 10, ("Hz", "y"): 11}
# the two field components each curl reads (the reference's keyword names)
_SOURCES = {"Hx": ("Ey", "Ez"), "Hy": ("Ex", "Ez"), "Hz": ("Ex", "Ey"),
            "Ex": ("Hy", "Hz"), "Ey": ("Hx", "Hz"), "Ez": ("Hx", "Hy")}
_GRID_SHAPE = {  # node counts from the updated component's shape
    "Ex": lambda s: (s[0] + 1, s[1], s[2]), "Ey": lambda s: (s[0], s[1] + 1, s[2]),
    "Ez": lambda s: (s[0], s[1], s[2] + 1), "Hx": lambda s: (s[0], s[1] + 1, s[2] + 1),
    "Hy": lambda s: (s[0] + 1, s[1], s[2] + 1), "Hz": lambda s: (s[0] + 1, s[1] + 1, s[2])}


def _lib():
    return nat.load_library()


def _check(rc):
    nat.check(rc, _lib())


def _f32(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_cuda:
        raise TypeError(f"{name} must be a float32 HIP tensor")
    return t.contiguous()


def _stream():
    return nat.current_stream()


class _Launch:
    """What ``NativeFDTDModule.<name>(**kwargs)`` returns (backend.py:66-85)."""

    def __init__(self, kernel: Callable[..., None], kwargs: dict):
        self.kernel = kernel
        self.kwargs = kwargs

    def launchRaw(self):  # noqa: N802 - the reference's entry point name
        self.kernel(**self.kwargs)
        return None


# ----------------------------------------------------------------- curl ops
_ZEROS: dict = {}


def _zeros(shape, device):
    """A cached zero tensor: stands in for the arrays an op's kernel may read but
    the op's result does not use (the other field components, the spacing and
    CPML profiles of axes the component's curl does not differentiate along)."""
    key = (tuple(shape), str(device))
    t = _ZEROS.get(key)
    if t is None:
        t = torch.zeros(shape, dtype=torch.float32, device=device)
        _ZEROS[key] = t
    return t


def _component_shape(c, nodes):
    nx, ny, nz = nodes
    return {"Ex": (nx - 1, ny, nz), "Ey": (nx, ny - 1, nz), "Ez": (nx, ny, nz - 1),
            "Hx": (nx, ny - 1, nz - 1), "Hy": (nx - 1, ny, nz - 1), "Hz": (nx - 1, ny - 1, nz)}[c]


def _curl_update(comp, kw, cpml):
    """One component's H or E update through mw_update_magnetic / _electric.

    ``cpml``: None (standard), "dense" (psi of the component's full shape) or
    "compressed" (slab psi with its low/high lengths)."""
    field = _f32(kw[comp], comp)
    dev = field.device
    shape = tuple(field.shape)
    nodes = _GRID_SHAPE[comp](shape)
    electric = comp[0] == "E"
    keep = [field]
    g = nat.MwGrid()
    g.nx, g.ny, g.nz = nodes
    for a in range(3):
        g.bc_lo[a] = g.bc_hi[a] = 0  # MW_BC_NONE
    given = {}
    for ax in _CURL_AXES[comp]:
        t = _f32(kw["invD" + ax], "invD" + ax)
        keep.append(t)
        given[ax] = nat.ptr(t)
    for a, ax in enumerate("xyz"):
        # spacing the update does not use: zeros of the axis' node count (long
        # enough for either family's array)
        filler = nat.ptr(_zeros((nodes[a],), dev))
        g.inv_e[a] = given.get(ax, filler) if electric else filler
        g.inv_h[a] = filler if electric else given.get(ax, filler)
    if electric:
        for ax in _CURL_AXES[comp]:
            a = _AXIS[ax]
            g.bc_lo[a] = int(kw[f"{ax}LowBoundaryMode"])
            g.bc_hi[a] = int(kw[f"{ax}HighBoundaryMode"])
    ci = _COMP.index(comp)
    cf = nat.MwCoeffs()
    for which, arr_slot, val_slot in (("Decay", cf.decay, cf.decay_value), ("Curl", cf.curl, cf.curl_value)):
        arr = kw.get(comp + which)
        uniform = kw.get("uniform" + which)
        if uniform is not None:
            arr_slot[ci] = None
            val_slot[ci] = float(uniform)
        else:
            t = _f32(arr, comp + which)
            keep.append(t)
            arr_slot[ci] = nat.ptr(t)
    f = nat.MwFields()
    srcs = [_f32(kw[s], s) for s in _SOURCES[comp]]
    keep += srcs
    # every array the kernel indexes must have the extent the grid implies
    for s_name, t in zip(_SOURCES[comp], srcs):
        if tuple(t.shape) != _component_shape(s_name, nodes):
            raise ValueError(f"{s_name} has shape {tuple(t.shape)}, expected {_component_shape(s_name, nodes)} "
                             f"for {comp} of shape {shape}")
    for which in ("Decay", "Curl"):
        t = kw.get(comp + which)
        if kw.get("uniform" + which) is None and tuple(t.shape) != shape:
            raise ValueError(f"{comp}{which} must have the shape of {comp}")
    for ax in _CURL_AXES[comp]:
        a = _AXIS[ax]
        need = nodes[a] if electric else nodes[a] - 1
        if kw["invD" + ax].numel() < need:
            raise ValueError(f"invD{ax} holds {kw['invD' + ax].numel()} values, the {comp} update needs {need}")
        if cpml is not None:
            tag = comp + ax.upper()
            for name in ("InvKappa" + tag, ("B" if electric else "By") + tag, ("C" if electric else "Cy") + tag):
                if kw[name].numel() < need:
                    raise ValueError(f"{name} holds {kw[name].numel()} values, the {comp} update needs {need}")
    ptrs = {comp: nat.ptr(field), _SOURCES[comp][0]: nat.ptr(srcs[0]), _SOURCES[comp][1]: nat.ptr(srcs[1])}
    for a in range(3):  # components the op does not pass: zero arrays of their shape
        for c, slot in ((_COMP[a], f.e), (_COMP[a + 3], f.h)):
            slot[a] = ptrs[c] if c in ptrs else nat.ptr(_zeros(_component_shape(c, nodes), dev))
    cp = nat.MwCpml()
    if cpml is not None:
        cp.enabled = 1
        prof = {}
        for ax in _CURL_AXES[comp]:
            tag = comp + ax.upper()
            # the reference names the H profiles ByHxY, CyHxY, ... and the E ones BExY, CExY, ...
            b_name, c_name = ("B" + tag, "C" + tag) if electric else ("By" + tag, "Cy" + tag)
            k = _f32(kw["InvKappa" + tag], "InvKappa" + tag)
            b = _f32(kw[b_name], b_name)
            c = _f32(kw[c_name], c_name)
            keep += [k, b, c]
            prof[ax] = (nat.ptr(k), nat.ptr(b), nat.ptr(c))
        for a, ax in enumerate("xyz"):
            z = nat.ptr(_zeros((nodes[a],), dev))
            own = prof.get(ax)
            k, b, c = own if own is not None else (z, z, z)
            for dst, v in ((cp.inv_kappa_e, k), (cp.b_e, b), (cp.c_e, c)):
                dst[a] = v if electric else z
            for dst, v in ((cp.inv_kappa_h, k), (cp.b_h, b), (cp.c_h, c)):
                dst[a] = z if electric else v
        for ax in _CURL_AXES[comp]:
            q = _PSI_SLOT[(comp, ax)]
            psi = _f32(kw["Psi" + comp + ax.upper()], "Psi" + comp + ax.upper())
            keep.append(psi)
            a = _AXIS[ax]
            size = shape[a]
            if cpml == "dense":
                if tuple(psi.shape) != shape:
                    raise ValueError(f"Psi{comp}{ax.upper()} must have the shape of {comp}")
                low, hs, hl = size, size, 0  # one slab spanning the axis: the dense array
            else:
                low = int(kw[f"psi{comp}{ax.upper()}LowLength"])
                hs = int(kw[f"psi{comp}{ax.upper()}HighStart"])
                hl = int(kw[f"psi{comp}{ax.upper()}HighLength"])
                want = list(shape)
                want[a] = low + hl
                if tuple(psi.shape) != tuple(want):
                    raise ValueError(f"Psi{comp}{ax.upper()} has shape {tuple(psi.shape)}, expected {tuple(want)} "
                                     f"(low + high = {low + hl} samples along {ax})")
            cp.psi[q] = nat.ptr(psi) if psi.numel() else None
            cp.psi_low[q], cp.psi_high_start[q], cp.psi_high_len[q] = low, hs, hl
    mask = 1 << (ci % 3)
    if electric:
        _check(_lib().mw_update_electric(ctypes.byref(g), ctypes.byref(cf), ctypes.byref(cp), None,
                                         ctypes.byref(f), mask, _stream()))
    else:
        _check(_lib().mw_update_magnetic(ctypes.byref(g), ctypes.byref(cf), ctypes.byref(cp), ctypes.byref(f), mask,
                                         _stream()))
    if field.data_ptr() != kw[comp].data_ptr():
        kw[comp].copy_(field)


def _curl_op(comp, cpml):
    def run(**kw):
        _curl_update(comp, kw, cpml)
    run.__name__ = f"update_{comp}_{cpml or 'standard'}"
    return run


# ----------------------------------------------------------------- sources
def _add_source_patch(*, field, sourcePatch, offsetI, offsetJ, offsetK, signal):
    f, p = _f32(field, "field"), _f32(sourcePatch, "sourcePatch")
    _check(_lib().mw_add_source_patch(nat.ptr(f), *f.shape, nat.ptr(p), *p.shape, int(offsetI), int(offsetJ),
                                      int(offsetK), float(signal), _stream()))


_TIME_KIND = {0: 0, 1: 1, 2: 2}  # cw, gaussian pulse, ricker (kernels/sources.cu:71-93)


def _add_time_shifted_source_patch(*, field, sourcePatch, delayPatch, activationDelayPatch, offsetI, offsetJ,
                                   offsetK, timeKind, time, frequency, fwidth, amplitude, phase, delay, causalGate):
    if int(causalGate) != 0:
        raise NotImplementedError("addTimeShiftedSourcePatch3D: the causal activation gate has no counterpart here")
    f, p = _f32(field, "field"), _f32(sourcePatch, "sourcePatch")
    d = delayPatch.to(dtype=torch.float64).contiguous()  # the entry point takes float64 delays
    _check(_lib().mw_add_time_shifted_patch(nat.ptr(f), *f.shape, nat.ptr(p), ctypes.c_void_p(d.data_ptr()),
                                            *p.shape, int(offsetI), int(offsetJ), int(offsetK), float(time),
                                            _TIME_KIND[int(timeKind)], float(frequency), float(fwidth),
                                            float(amplitude), float(phase), float(delay), _stream()))


# ----------------------------------------------------------- TFSF aux line
def _update_auxiliary_magnetic(*, Magnetic, Electric, MagneticDecay, MagneticCurl):
    h, e = _f32(Magnetic, "Magnetic"), _f32(Electric, "Electric")
    d, c = _f32(MagneticDecay, "MagneticDecay"), _f32(MagneticCurl, "MagneticCurl")
    if e.numel() != h.numel() + 1 or d.numel() < h.numel() or c.numel() < h.numel():
        raise ValueError("the auxiliary line's H has one sample fewer than its E")
    _check(_lib().mw_update_aux_magnetic(nat.ptr(h), nat.ptr(e), nat.ptr(d), nat.ptr(c), h.numel(), _stream()))


def _update_auxiliary_electric(*, Electric, Magnetic, ElectricDecay, ElectricCurl, sourceIndex, sourceValue):
    e, h = _f32(Electric, "Electric"), _f32(Magnetic, "Magnetic")
    d, c = _f32(ElectricDecay, "ElectricDecay"), _f32(ElectricCurl, "ElectricCurl")
    if e.numel() != h.numel() + 1 or d.numel() < e.numel() or c.numel() < e.numel():
        raise ValueError("the auxiliary line's H has one sample fewer than its E")
    _check(_lib().mw_update_aux_electric(nat.ptr(e), nat.ptr(h), nat.ptr(d), nat.ptr(c), e.numel(), int(sourceIndex),
                                         float(sourceValue), _stream()))


# --------------------------------------------------------------- monitors
def _accumulate_dft_batched(*, Ex, Ey, Ez, ExRealAccum, ExImagAccum, EyRealAccum, EyImagAccum, EzRealAccum,
                            EzImagAccum, weightedCos, weightedSin):
    wc, ws = _f32(weightedCos, "weightedCos"), _f32(weightedSin, "weightedSin")
    if wc.numel() != ws.numel():
        raise ValueError("weightedCos and weightedSin must have one value per frequency")
    for f, re, im in ((Ex, ExRealAccum, ExImagAccum), (Ey, EyRealAccum, EyImagAccum), (Ez, EzRealAccum, EzImagAccum)):
        f, re, im = _f32(f, "field"), _f32(re, "realAccum"), _f32(im, "imagAccum")
        if re.numel() != wc.numel() * f.numel() or im.numel() != re.numel():
            raise ValueError("DFT accumulators must have shape (frequencies,) + field shape")
        _check(_lib().mw_accumulate_dft(nat.ptr(f), f.numel(), wc.numel(), nat.ptr(wc), nat.ptr(ws), nat.ptr(re),
                                        nat.ptr(im), _stream()))


def _weights(cos, sin, device):
    return torch.tensor([float(cos)], dtype=torch.float32, device=device), \
        torch.tensor([float(sin)], dtype=torch.float32, device=device)


def _accumulate_point_observers(*, field, pointI, pointJ, pointK, realAccum, imagAccum, weightedCos, weightedSin):
    f = _f32(field, "field")
    sx, sy, sz = f.shape
    idx = [p.to(device=f.device, dtype=torch.int64).reshape(-1) for p in (pointI, pointJ, pointK)]
    for p, n, name in zip(idx, (sx, sy, sz), ("pointI", "pointJ", "pointK")):
        if p.numel() and (int(p.min()) < 0 or int(p.max()) >= n):
            raise ValueError(f"{name} outside the field's extent {n}")
    lin = ((idx[0] * sy + idx[1]) * sz + idx[2]).to(torch.int32).contiguous()
    if realAccum.numel() != lin.numel() or imagAccum.numel() != lin.numel():
        raise ValueError("point accumulators must hold one value per point")
    wc, ws = _weights(weightedCos, weightedSin, f.device)
    re, im = _f32(realAccum, "realAccum"), _f32(imagAccum, "imagAccum")
    _check(_lib().mw_accumulate_points(nat.ptr(f), ctypes.c_void_p(lin.data_ptr()), lin.numel(), 1, nat.ptr(wc),
                                       nat.ptr(ws), nat.ptr(re), nat.ptr(im), _stream()))


def _accumulate_plane_observer(*, field, planeRealAccum, planeImagAccum, axisCode, planeIndex, weightedCos,
                               weightedSin):
    f = _f32(field, "field")
    wc, ws = _weights(weightedCos, weightedSin, f.device)
    re, im = _f32(planeRealAccum, "planeRealAccum"), _f32(planeImagAccum, "planeImagAccum")
    cells = f.numel() // f.shape[int(axisCode)]
    if re.numel() != cells or im.numel() != cells:
        raise ValueError(f"plane accumulators must hold the plane's {cells} samples")
    _check(_lib().mw_accumulate_plane(nat.ptr(f), *f.shape, int(axisCode), int(planeIndex), 1, nat.ptr(wc),
                                      nat.ptr(ws), nat.ptr(re), nat.ptr(im), _stream()))


# --------------------------------------------------------------- dispersive
def _pole(kind, e, drive, pol, cur, decay, restoring, dt):
    e, drive, cur = _f32(e, "ElectricField"), _f32(drive, "Drive"), _f32(cur, "PolarizationCurrent")
    pol = _f32(pol, "Polarization") if pol is not None else None
    for t in (drive, cur) + ((pol,) if pol is not None else ()):
        if t.numel() != e.numel():
            raise ValueError("pole arrays must have the electric field's shape")
    polp = nat.ptr(pol) if pol is not None else None
    _check(_lib().mw_update_pole_current(int(kind), nat.ptr(e), nat.ptr(drive), polp, nat.ptr(cur), e.numel(),
                                         float(decay), float(restoring), float(dt), _stream()))


def _update_debye_current(*, ElectricField, Polarization, PolarizationCurrent, DebyeDrive, decay, dt):
    _pole(0, ElectricField, DebyeDrive, Polarization, PolarizationCurrent, decay, 0.0, dt)


def _update_drude_current(*, ElectricField, PolarizationCurrent, DrudeDrive, decay):
    _pole(1, ElectricField, DrudeDrive, None, PolarizationCurrent, decay, 0.0, 1.0)


def _update_lorentz_current(*, ElectricField, Polarization, PolarizationCurrent, LorentzDrive, decay, restoring, dt):
    _pole(2, ElectricField, LorentzDrive, Polarization, PolarizationCurrent, decay, restoring, dt)


def _apply_polarization_current(*, ElectricField, PolarizationCurrent, InvPermittivity, dt):
    e = _f32(ElectricField, "ElectricField")
    cur, inv = _f32(PolarizationCurrent, "PolarizationCurrent"), _f32(InvPermittivity, "InvPermittivity")
    if cur.numel() != e.numel() or inv.numel() != e.numel():
        raise ValueError("PolarizationCurrent and InvPermittivity must have the electric field's shape")
    _check(_lib().mw_apply_polarization_current(nat.ptr(e), nat.ptr(cur), nat.ptr(inv), e.numel(), float(dt),
                                                _stream()))


_KERNELS: dict[str, Callable[..., None]] = {}
for _c in _COMP:
    _fam = "Magnetic" if _c[0] == "H" else "Electric"
    _KERNELS[f"update{_fam}Field{_c}Standard3D"] = _curl_op(_c, None)
    _KERNELS[f"update{_fam}Field{_c}CpmlCompressed3D"] = _curl_op(_c, "compressed")
    _KERNELS[f"update{_fam}Field{_c}{'3D' if _c[0] == 'H' else 'Cpml3D'}"] = _curl_op(_c, "dense")
_KERNELS.update({
    "addSourcePatch3D": _add_source_patch,
    "addTimeShiftedSourcePatch3D": _add_time_shifted_source_patch,
    "updateAuxiliaryMagnetic1D": _update_auxiliary_magnetic,
    "updateAuxiliaryElectric1D": _update_auxiliary_electric,
    "accumulateRunningDftYee3DBatched": _accumulate_dft_batched,
    "accumulatePointObservers3D": _accumulate_point_observers,
    "accumulatePlaneObserver3D": _accumulate_plane_observer,
    "updateDebyeCurrent3D": _update_debye_current,
    "updateDrudeCurrent3D": _update_drude_current,
    "updateLorentzCurrent3D": _update_lorentz_current,
    "applyPolarizationCurrent3D": _apply_polarization_current,
})


class NativeFDTDModule:
    """Attribute access by kernel name returns a binder; ``binder(**kwargs)``
    returns a launch whose ``launchRaw()`` runs the op (backend.py:89-104)."""

    def __getattr__(self, name: str):
        kernel = _KERNELS.get(name)
        if kernel is None:
            raise AttributeError(f"Native HIP FDTD backend does not implement kernel {name!r}.")

        def bind(**kwargs):
            return _Launch(kernel, kwargs)

        bind.__name__ = name
        object.__setattr__(self, name, bind)
        return bind


_NATIVE_MODULE = NativeFDTDModule()


def get_native_fdtd_module() -> NativeFDTDModule:
    """The module instance (backend.py:110-115); loads the HIP library and fails
    loudly when it is not built or no HIP device is present."""
    if not torch.cuda.is_available():
        raise RuntimeError("Native HIP FDTD backend requires a HIP device (torch.cuda.is_available() is False).")
    nat.load_library()
    return _NATIVE_MODULE


__all__ = ["NativeFDTDModule", "get_native_fdtd_module"]
