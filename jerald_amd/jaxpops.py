"""Drop-in replacement for the reference's ``jaxpops_bind`` operators on one
MI355X: same names, argument layout, output shapes and normalisations
(jaxpops_bind.py:4), with torch tensors on the GPU instead of JAX arrays.

Each differentiable operator is a ``torch.autograd.Function`` whose backward
calls hand-written adjoint kernels of libldl:

=============  ==============================  ================================
operator       reference                       backward
=============  ==============================  ================================
ppaint         jaxpops_bind.py:136-249         readout-gradient * mass, readout
preadout       jaxpops_bind.py:255-365         readout-gradient * v, paint(v)
ppaint3D       jaxpops_bind.py:371-465         readout-gradient3, readout3
               (the reference has no VJP;      (an extension)
               this one does)
preadout3D     jaxpops_bind.py:471-590         readout-gradient3, paint3D(v);
                                               vjp=True has no VJP (as reference)
pfft           jaxpops_bind.py:647-703         rule-dependent, see below
pifft          jaxpops_bind.py:709-761         rule-dependent, see below
c2f            model.py:11-31                  rule-dependent, see below
=============  ==============================  ================================

FFT VJP rules.  torch's gradient of a complex tensor is the complex conjugate
of JAX's cotangent.  ``set_vjp_rule("reference")`` (default) reproduces the
reference's custom rules translated to torch's convention:
pfft: g -> pifft(g) (jaxpops_bind.py:699-701), pifft: g -> pfft(g)
(:758-759), c2f doubles the interior k_z cotangent (model.py:23-29).  These
compose to the exact gradient wherever the reference relies on them (r2c
followed by c2r, and c2f on a real kernel), but alone they are not the exact
adjoints.  ``set_vjp_rule("exact")`` gives the exact adjoints (pfft: g ->
pifft(g / c), pifft: g -> c pfft(g), c2f identity) with c = 1 on the k_z = 0
and N/2 planes and 2 elsewhere.

Single-GPU contract: ``comm`` must be None or a communicator of size 1 and
``edges`` the full mesh; the MPI layout primitives are no-ops.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L

__all__ = ["initLayout", "clean", "ppaint", "preadout", "ppaint3D", "preadout3D", "buildplan", "pfft", "pifft",
           "c2f", "local_positions", "set_vjp_rule", "get_vjp_rule", "set_paint_mode"]

_VJP_RULE = "reference"
_PAINT_FLAGS = L.LDL_PAINT_FIXED


def set_vjp_rule(rule: str):
    """'reference' (the reference's pfft/pifft/c2f VJP rules) or 'exact'."""
    global _VJP_RULE
    if rule not in ("reference", "exact"):
        raise ValueError("rule must be 'reference' or 'exact'")
    _VJP_RULE = rule


def get_vjp_rule() -> str:
    return _VJP_RULE


def set_paint_mode(mode: str):
    """'fixed' (64-bit fixed point, tile-binned, deterministic), 'direct'
    (the same fixed point with one global atomic per corner) or 'float'
    (floating-point atomics)."""
    global _PAINT_FLAGS
    _PAINT_FLAGS = {"fixed": L.LDL_PAINT_FIXED, "direct": L.LDL_PAINT_FIXED | L.LDL_PAINT_DIRECT,
                    "float": L.LDL_PAINT_FLOAT_ATOMIC}[mode]


# ------------------------------------------------------------ validation
def _comm_ok(comm):
    if comm is None:
        return
    size = getattr(comm, "size", None)
    if size is None and hasattr(comm, "Get_size"):
        size = comm.Get_size()
    if size not in (None, 1):
        raise NotImplementedError("single-GPU build: comm must be None or of size 1")


def _nmesh(Nmesh):
    n = np.asarray(Nmesh.cpu() if torch.is_tensor(Nmesh) else Nmesh).reshape(-1)
    if n.size == 1:
        n = np.repeat(n, 3)
    return tuple(int(v) for v in n)


def _box(BoxSize):
    b = np.asarray(BoxSize.cpu() if torch.is_tensor(BoxSize) else BoxSize, dtype=np.float64).reshape(-1)
    if b.size == 1:
        b = np.repeat(b, 3)
    return tuple(float(v) for v in b)


def _check_edges(edges, nm):
    if edges is None:
        return
    e = np.asarray(edges.cpu() if torch.is_tensor(edges) else edges)
    if e.shape != (3, 1, 2) or any(int(e[d, 0, 0]) != 0 or int(e[d, 0, 1]) != nm[d] for d in range(3)):
        # pmesh.py:62 sets edges[1,:,1] = Nmesh[2]; accept it for cubic meshes only
        raise NotImplementedError("single-GPU build: edges must be [[[0, N0]], [[0, N1]], [[0, N2]]]")


def _lyidx(lyidx):
    if lyidx > 4:
        raise Exception("lyidx needs to be smaller than 5")  # jaxpops_bind.py:157-158


def _dims(dims, nm):
    if dims is not None and tuple(int(d) for d in dims) != nm:
        raise NotImplementedError("single-GPU build: dims must equal Nmesh")


def _pos(pos):
    L.require_gpu(pos, "pos")
    if pos.dim() != 2 or pos.shape[1] != 3:
        raise ValueError("pos must be (Np, 3)")
    return pos.contiguous()


# ----------------------------------------------------------- raw kernels
def _paint_raw(pos, mass, nm, box, out_scale=1.0):
    """mass: python scalar, (Np,) tensor, or (Np,3) tensor (3-component)."""
    dt = pos.dtype
    dev = pos.device
    pd = L.pos_desc(box)
    st = L.stream_ptr(dev)
    if torch.is_tensor(mass) and mass.dim() == 2:
        out = torch.empty((3,) + nm, dtype=dt, device=dev)
        flags = _PAINT_FLAGS
        nb = L.load().ldl_paint_workspace_bytes(L.i64x3(nm), 3, pos.shape[0], flags)
        ws = L.workspaces.get(dev, "paint", nb)
        m = mass.to(dt).contiguous()
        L.call(f"ldl_ppaint3D_{L.sfx(dt)}", L.ptr(pos), pd, pos.shape[0], L.ptr(m), L.i64x3(nm), float(out_scale),
               flags, L.ptr(out), L.ptr(ws), ws.numel(), st)
        return out
    out = torch.empty(nm, dtype=dt, device=dev)
    flags = _PAINT_FLAGS
    nb = L.load().ldl_paint_workspace_bytes(L.i64x3(nm), 1, pos.shape[0], flags)
    ws = L.workspaces.get(dev, "paint1", nb)
    if torch.is_tensor(mass) and mass.dim() >= 1 and mass.numel() == pos.shape[0]:
        m = mass.to(dt).reshape(-1).contiguous()
        L.call(f"ldl_ppaint_{L.sfx(dt)}", L.ptr(pos), pd, pos.shape[0], L.ptr(m), 0.0, L.i64x3(nm), float(out_scale),
               flags, L.ptr(out), L.ptr(ws), ws.numel(), st)
    else:
        ms = float(mass.item() if torch.is_tensor(mass) else mass)
        L.call(f"ldl_ppaint_{L.sfx(dt)}", L.ptr(pos), pd, pos.shape[0], None, ms, L.i64x3(nm), float(out_scale),
               flags, L.ptr(out), L.ptr(ws), ws.numel(), st)
    return out


def _readout_raw(pos, field, nm, box, vjpdim=-1):
    out = torch.empty(pos.shape[0], dtype=pos.dtype, device=pos.device)
    f = field.to(pos.dtype).contiguous()
    L.call(f"ldl_preadout_{L.sfx(pos.dtype)}", L.ptr(pos), L.pos_desc(box), pos.shape[0], L.ptr(f), L.i64x3(nm),
           int(vjpdim), L.ptr(out), L.stream_ptr(pos.device))
    return out


def _readout3_raw(pos, field, nm, box, vjp, pscale=None):
    out = torch.empty((pos.shape[0], 3), dtype=pos.dtype, device=pos.device)
    f = field.to(pos.dtype).contiguous()
    ps = None if pscale is None else pscale.to(pos.dtype).reshape(-1).contiguous()
    L.call(f"ldl_preadout3D_{L.sfx(pos.dtype)}", L.ptr(pos), L.pos_desc(box), pos.shape[0], L.ptr(f), L.i64x3(nm),
           int(bool(vjp)), L.ptr(ps), L.ptr(out), L.stream_ptr(pos.device))
    return out


def _readout_grad3_raw(pos, fields3, v3, nm, box):
    out = torch.empty((pos.shape[0], 3), dtype=pos.dtype, device=pos.device)
    f = fields3.to(pos.dtype).contiguous()
    v = v3.to(pos.dtype).contiguous()
    L.call(f"ldl_readout_grad3_{L.sfx(pos.dtype)}", L.ptr(pos), L.pos_desc(box), pos.shape[0], L.ptr(f), L.ptr(v),
           L.i64x3(nm), L.ptr(out), L.stream_ptr(pos.device))
    return out


def _pfft_raw(x, scale=1.0):
    L.require_gpu(x, "data")
    x = x.contiguous()
    nm = tuple(x.shape)
    out = torch.empty(nm[:2] + (nm[2] // 2 + 1, 2), dtype=x.dtype, device=x.device)
    L.call(f"ldl_pfft_{L.sfx(x.dtype)}", L.ptr(x), L.ptr(out), L.i64x3(nm), float(scale), L.stream_ptr(x.device))
    return torch.view_as_complex(out)


def _pifft_raw(X, nm, scale=1.0):
    L.require_gpu(X, "data")
    rdt = torch.float32 if X.dtype == torch.complex64 else torch.float64
    Xr = torch.view_as_real(X.contiguous()) if X.is_complex() else X.contiguous()
    out = torch.empty(nm, dtype=rdt, device=X.device)
    nb = L.load().ldl_fft_workspace_bytes(L.i64x3(nm), int(rdt == torch.float64))
    ws = L.workspaces.get(X.device, "fft", nb)
    L.call(f"ldl_pifft_{L.sfx(rdt)}", L.ptr(Xr), L.ptr(out), L.i64x3(nm), float(scale), L.ptr(ws), ws.numel(),
           L.stream_ptr(X.device))
    return out


def hermitian_weight(n2, device, dtype):
    """c(k_z): 1 on the k_z = 0 and N/2 planes, 2 elsewhere (a table for
    callers such as Parseval checks; the VJPs apply it in ``_zweight``)."""
    c = torch.full((n2 // 2 + 1,), 2.0, dtype=dtype, device=device)
    c[0] = 1.0
    c[-1] = 1.0
    return c


ZW_ONE, ZW_C, ZW_INV_C, ZW_C2F, ZW_REAL = 0, 1, 2, 3, 4


def _zweight(v, nm, mode, scale=1.0):
    """New tensor w(k_z) * scale * v over a half spectrum (ldl_spectrum_zweight):
    mode ZW_C multiplies by c, ZW_INV_C divides by it, ZW_C2F doubles z bins
    [1, int(N1/2)); ``v`` complex, or real for a real kernel's bins."""
    L.require_gpu(v, "spectrum")
    v = v.resolve_conj().contiguous()
    out = torch.empty_like(v)
    if v.is_complex():
        rdt = torch.float32 if v.dtype == torch.complex64 else torch.float64
        a, b = torch.view_as_real(v), torch.view_as_real(out)
    else:
        rdt, a, b = v.dtype, v, out
        mode |= ZW_REAL
    L.call(f"ldl_spectrum_zweight_{L.sfx(rdt)}", L.ptr(a), L.ptr(b), L.i64x3(nm), int(mode), float(scale),
           L.stream_ptr(v.device))
    return out


# ------------------------------------------------------- autograd ops
class _PPaint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, mass, nm, box):
        ctx.nm, ctx.box = nm, box
        ctx.scalar_mass = not torch.is_tensor(mass)
        if ctx.scalar_mass:
            ctx.mass_value = float(mass)
            ctx.save_for_backward(pos)
        else:
            ctx.save_for_backward(pos, mass)
        return _paint_raw(pos, mass, nm, box)

    @staticmethod
    def backward(ctx, v):
        v = v.contiguous()
        if ctx.scalar_mass:
            (pos,) = ctx.saved_tensors
            gpos = _readout3_raw(pos, v, ctx.nm, ctx.box, True) * ctx.mass_value if ctx.needs_input_grad[0] else None
            return gpos, None, None, None
        pos, mass = ctx.saved_tensors
        gpos = _readout3_raw(pos, v, ctx.nm, ctx.box, True, pscale=mass) if ctx.needs_input_grad[0] else None
        gmass = _readout_raw(pos, v, ctx.nm, ctx.box).reshape(mass.shape).to(mass.dtype) if ctx.needs_input_grad[1] else None
        return gpos, gmass, None, None


class _PReadout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, field, nm, box, vjpdim):
        ctx.nm, ctx.box, ctx.vjpdim = nm, box, vjpdim
        ctx.save_for_backward(pos, field)
        return _readout_raw(pos, field, nm, box, vjpdim)

    @staticmethod
    def backward(ctx, vmass):
        pos, field = ctx.saved_tensors
        vmass = vmass.contiguous()
        if ctx.vjpdim >= 0 and ctx.needs_input_grad[0]:
            raise NotImplementedError("second derivative of the CIC readout is not defined")
        gpos = _readout3_raw(pos, field, ctx.nm, ctx.box, True, pscale=vmass) if ctx.needs_input_grad[0] else None
        gfield = None
        if ctx.needs_input_grad[1]:
            if ctx.vjpdim >= 0:
                raise NotImplementedError("vjpdim readout is not differentiable w.r.t. the field")
            gfield = _paint_raw(pos, vmass, ctx.nm, ctx.box).reshape(field.shape)
        return gpos, gfield, None, None, None


class _PPaint3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, mass3, nm, box):
        ctx.nm, ctx.box = nm, box
        ctx.save_for_backward(pos, mass3)
        return _paint_raw(pos, mass3, nm, box)

    @staticmethod
    def backward(ctx, v):
        pos, mass3 = ctx.saved_tensors
        v = v.contiguous()
        gpos = _readout_grad3_raw(pos, v, mass3, ctx.nm, ctx.box) if ctx.needs_input_grad[0] else None
        gmass = _readout3_raw(pos, v, ctx.nm, ctx.box, False) if ctx.needs_input_grad[1] else None
        return gpos, gmass, None, None


class _PReadout3D(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, field, nm, box, vjp):
        ctx.nm, ctx.box, ctx.vjp = nm, box, vjp
        ctx.save_for_backward(pos, field)
        return _readout3_raw(pos, field, nm, box, vjp)

    @staticmethod
    def backward(ctx, vmass):
        if ctx.vjp:
            raise Exception("custom_vjp not implemented for 3D readout with vjp = True")  # jaxpops_bind.py:575-576
        pos, field = ctx.saved_tensors
        vmass = vmass.contiguous()
        gpos = _readout_grad3_raw(pos, field, vmass, ctx.nm, ctx.box) if ctx.needs_input_grad[0] else None
        gfield = _paint_raw(pos, vmass, ctx.nm, ctx.box) if ctx.needs_input_grad[1] else None
        return gpos, gfield, None, None, None


class _PFFT(torch.autograd.Function):
    """scale * r2c; scale 1 is pfft, 1/N^3 PMesh.r2c (pmesh.py:112)."""

    @staticmethod
    def forward(ctx, data, rule, scale):
        ctx.rule, ctx.nm, ctx.scale = rule, tuple(data.shape), scale
        return _pfft_raw(data, scale)

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous()
        if ctx.rule == "exact":
            g = _zweight(g, ctx.nm, ZW_INV_C)
        return _pifft_raw(g, ctx.nm, ctx.scale), None, None


class _PIFFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, nm, rule):
        ctx.rule, ctx.nm = rule, nm
        return _pifft_raw(data, nm)

    @staticmethod
    def backward(ctx, g):
        G = _pfft_raw(g.contiguous())
        if ctx.rule == "exact":
            G = _zweight(G, ctx.nm, ZW_C)
        return G, None, None


class _C2F(torch.autograd.Function):
    @staticmethod
    def forward(ctx, obj, rule):
        ctx.rule = rule
        return obj.clone()

    @staticmethod
    def backward(ctx, v):
        if ctx.rule == "exact":
            return v, None
        # model.py:26-28: z bins [1, int(v.shape[1]/2)) get v + conj(v); the
        # second axis length is v.shape[1], the z length of the half spectrum
        # is inferred as 2 (nh - 1) (only the bin count matters for this mode)
        d0, d1, nh = v.shape
        return _zweight(v.resolve_conj(), (d0, d1, 2 * (nh - 1)), ZW_C2F), None


# --------------------------------------------------------- public API
def initLayout(commsize, lyidx=0):
    """No-op on one GPU (jaxpops_bind.py:57); commsize must be 1."""
    if int(commsize) != 1:
        raise NotImplementedError("single-GPU build: commsize must be 1")
    _lyidx(lyidx)
    return 0


def clean(lyidx=0):
    """No-op on one GPU (jaxpops_bind.py:82); the ops keep no layout state."""
    _lyidx(lyidx)
    return 0


def buildplan(Nmesh, comm=None):
    """Validates the mesh and prepares twiddles; returns (localL, localS) =
    (N0, 0) as an int32 tensor (jaxpops_bind.py:597-606)."""
    _comm_ok(comm)
    nm = _nmesh(Nmesh)
    out = (ctypes.c_int32 * 2)()
    L.call("ldl_buildplan", L.i64x3(nm), out)
    return torch.tensor([out[0], out[1]], dtype=torch.int32)


def ppaint(pos, mass, Nmesh, BoxSize, edges=None, dims=None, comm=None, lyidx=0, bwd=False):
    """CIC paint (jaxpops_bind.py:136-178): output (d0, d1, d2) = deposit of
    ``mass`` (scalar or (Np,)) at ``pos`` (Np, 3), positions in box units."""
    _comm_ok(comm)
    _lyidx(lyidx)
    nm = _nmesh(Nmesh)
    _check_edges(edges, nm)
    _dims(dims, nm)
    pos = _pos(pos)
    if torch.is_tensor(mass) and mass.dim() >= 1:
        mass = torch.broadcast_to(mass, (pos.shape[0],)).to(pos.device)
    elif torch.is_tensor(mass):
        mass = float(mass.item())
    return _PPaint.apply(pos, mass, nm, _box(BoxSize))


def preadout(pos, field, Nmesh, BoxSize, edges=None, dims=None, comm=None, lyidx=0, vjpdim=-1):
    """CIC readout (jaxpops_bind.py:255-294) -> (Np,); vjpdim in {0,1,2}
    returns the derivative along that axis (box units)."""
    _comm_ok(comm)
    _lyidx(lyidx)
    if vjpdim > 2:
        raise Exception("vjpdim must be between 0 and 2")  # jaxpops_bind.py:277-278
    nm = _nmesh(Nmesh)
    _check_edges(edges, nm)
    pos = _pos(pos)
    field = field.reshape(nm)
    return _PReadout.apply(pos, field, nm, _box(BoxSize), int(vjpdim))


def ppaint3D(pos, mass, Nmesh, BoxSize, edges=None, dims=None, comm=None, lyidx=0):
    """Three-component paint (jaxpops_bind.py:371-414): mass (Np, 3) ->
    (3, d0, d1, d2)."""
    _comm_ok(comm)
    _lyidx(lyidx)
    if mass.shape[1] != 3:
        raise Exception("Invalid input mass shape")  # jaxpops_bind.py:393-394
    nm = _nmesh(Nmesh)
    _check_edges(edges, nm)
    pos = _pos(pos)
    return _PPaint3D.apply(pos, mass.contiguous(), nm, _box(BoxSize))


def preadout3D(pos, field, Nmesh, BoxSize, edges=None, dims=None, comm=None, lyidx=0, vjp=False):
    """3D readout (jaxpops_bind.py:471-516): vjp=False reads 3 fields
    (3, d0, d1, d2); vjp=True returns the gradient of one field.  -> (Np, 3)."""
    _comm_ok(comm)
    _lyidx(lyidx)
    if not vjp and len(field) != 3:
        raise Exception("Invalid input field shape")  # jaxpops_bind.py:496-497
    nm = _nmesh(Nmesh)
    _check_edges(edges, nm)
    pos = _pos(pos)
    field = field.reshape(nm if vjp else (3,) + nm)
    return _PReadout3D.apply(pos, field, nm, _box(BoxSize), bool(vjp))


def pfft(data, dims=None, scale=1.0):
    """Unnormalised r2c (jaxpops_bind.py:647-656): real (d0, d1, d2) ->
    complex (d0, d1, d2//2+1), times ``scale`` inside the transform (1/N^3 is
    PMesh.r2c, pmesh.py:112)."""
    if dims is not None:
        data = data.reshape(tuple(int(d) for d in dims))
    if data.dtype not in (torch.float32, torch.float64):
        raise NotImplementedError(f"Unsupported dtype {data.dtype}")
    return _PFFT.apply(data, _VJP_RULE, float(scale))


def pifft(data, dims):
    """Unnormalised c2r (jaxpops_bind.py:709-717): complex (d0, d1, d2//2+1)
    -> real (d0, d1, d2); imaginary parts of the k_z = 0, N/2 planes ignored."""
    nm = tuple(int(d) for d in dims)
    if data.dtype not in (torch.complex64, torch.complex128):
        data = data.to(torch.complex128 if data.dtype == torch.float64 else torch.complex64)
    return _PIFFT.apply(data, nm, _VJP_RULE)


def c2f(obj):
    """Identity whose backward follows the active VJP rule (model.py:11-31)."""
    return _C2F.apply(obj, _VJP_RULE)


def local_positions(pos, Nout, Nmesh, BoxSize, edges=None, commsize=1, commrank=0):
    """One rank owns every particle: returns the wrapped positions
    ``(pos % L)/(L/N) * (L/N)`` in the input order (jaxpops_bind.py:767-807)."""
    if int(commsize) != 1:
        raise NotImplementedError("single-GPU build: commsize must be 1")
    n = np.asarray(Nout).reshape(-1)
    if int(n.sum()) != pos.shape[0]:
        raise Exception(f"The sum of local sizes should match the total number of particles, got {n.sum()} and {len(pos)} instead.")
    box = torch.tensor(_box(BoxSize), dtype=pos.dtype, device=pos.device)
    cell = box / torch.tensor(_nmesh(Nmesh), dtype=pos.dtype, device=pos.device)
    return torch.remainder(pos, box) / cell * cell
