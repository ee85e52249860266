"""Learned diagonal preconditioner for CG on SENSE-NUFFT normal equations.

``SenseNormal`` is A = S^H F^H F S + lam I for coil maps S and the NUFFT F
of src/torchlinops/linops/nufft.py:123-184 (apodize, centred pad, centred
FFT, interpolate, scale; adjoint in reverse), with the interpolation on the
plan-driven kernels (the trajectory tensor is fixed, so its plan is built
once).

``PrecondNet`` is a 4-layer, 16-channel 3x3x3 convolutional network whose
convolutions run in BF16 on the matrix cores with FP32 accumulation
(``tnl_conv3d_fwd``; data gradient = the same kernel with flipped,
transposed weights; weight gradient ``tnl_conv3d_wgrad``).  It predicts a
positive diagonal image-space preconditioner M = diag(d) as a learned
correction of the analytic Jacobi diagonal: d = J exp(o), J = 1 /
(psf(0) sum_c |S_c|^2 + lam) (psf(0) = the constant diagonal of F^H F,
measured once with a unit impulse), o the network output over the features
of ``precond_inputs`` (how much of each voxel's diagonal is signal rather
than regularisation, and log J).  The last layer starts at zero, so the
untrained network is exactly Jacobi.  It is trained on solves that plain CG
generates: for each training problem A and right-hand sides z (A applied to
smooth random fields), plain CG provides x* = A^-1 z; the loss is the
relative error ||x_k(d) - x*||^2 / ||x*||^2 of k preconditioned-CG
iterations from zero, differentiated through the unrolled iterations (the
interpolation kernels' autograd pair supplies the operator's backward).

``solve`` runs preconditioned CG (``DeviceCG``) and keeps plain CG as the
judge: with ``judge_iters`` it first runs the preconditioned CG for that
many iterations and plain CG for the same time (the same count unless
applying the preconditioner costs time of its own) and keeps the
preconditioner only if its iterate is not behind plain CG's at that
matched cost, scored by CG's own objective (the A-norm error, from one
A x each; the recomputed residuals are reported beside it) - the loser's
work is counted, the winner continues its own Krylov sequence without a
restart; and if the preconditioned run stalls,
breaks down or ends above the tolerance on its recomputed residual, plain
CG continues from its iterate.  No counterpart in the
reference, whose CG (pcg.py) is unpreconditioned.
"""

from __future__ import annotations

import math
import time
from typing import Optional, Sequence

import torch
import torch.nn as nn
from torch.autograd import Function

from .._lib import check, lib, require_device, stream_of
from ..functional._interp import interpolate, interpolate_adjoint
from ..linops.nufft import NUFFT, shifted_layout
from ..linops.toeplitz import crop_wrapped, pad_wrapped
from .cg import DeviceCG, inner

__all__ = ["SenseNormal", "PrecondNet", "conv3d", "to_channel_last", "pack_conv_weights",
           "pack_conv_weights_transposed", "coil_features", "precond_inputs",
           "learned_diagonal", "pcg_unrolled", "train_preconditioner", "judge", "solve"]

HIDDEN = 16


def _kp(c: int) -> int:
    return 32 * math.ceil(27 * c / 32)


def _cpad(c: int) -> int:
    return 8 * math.ceil(c / 8)


def pack_conv_weights(w: torch.Tensor, cin_p: int) -> torch.Tensor:
    """fp32 [cout, cin, 3, 3, 3] -> bf16 [16, Kp], column tap * cin_p + ci."""
    cout, cin = w.shape[:2]
    W = torch.zeros(16, 27, cin_p, device=w.device, dtype=torch.float32)
    W[:cout, :, :cin] = w.detach().reshape(cout, cin, 27).permute(0, 2, 1)
    out = torch.zeros(16, _kp(cin_p), device=w.device, dtype=torch.float32)
    out[:, :27 * cin_p] = W.reshape(16, 27 * cin_p)
    return out.to(torch.bfloat16).contiguous()


def pack_conv_weights_transposed(w: torch.Tensor, cout_p: int) -> torch.Tensor:
    """Weights of the data gradient: rows = input channels, taps flipped
    (offset -d), columns tap * cout_p + co."""
    cout, cin = w.shape[:2]
    wf = w.detach().reshape(cout, cin, 27).flip(-1)
    W = torch.zeros(16, 27, cout_p, device=w.device, dtype=torch.float32)
    W[:cin, :, :cout] = wf.permute(1, 2, 0)
    out = torch.zeros(16, _kp(cout_p), device=w.device, dtype=torch.float32)
    out[:, :27 * cout_p] = W.reshape(16, 27 * cout_p)
    return out.to(torch.bfloat16).contiguous()


def to_channel_last(x: torch.Tensor, c_p: Optional[int] = None) -> torch.Tensor:
    """fp32 [B, C, Z, Y, X] -> bf16 [B, Z, Y, X, c_p] (zero padding channels)."""
    B, C = x.shape[:2]
    c_p = c_p or _cpad(C)
    out = torch.zeros(B, *x.shape[2:], c_p, device=x.device, dtype=torch.bfloat16)
    out[..., :C] = x.permute(0, 2, 3, 4, 1)
    return out


def _fwd(x: torch.Tensor, wpack: torch.Tensor, bias, cout: int, cout_p: int, relu: bool,
         out_f32: bool = False) -> torch.Tensor:
    B, Z, Y, X, cin_p = x.shape
    dt = torch.float32 if out_f32 else torch.bfloat16
    out = torch.zeros(B, Z, Y, X, cout_p, device=x.device, dtype=dt)
    with torch.cuda.device(x.device):
        check(lib().tnl_conv3d_fwd(
            x.data_ptr(), wpack.data_ptr(), None if bias is None else bias.data_ptr(),
            None if out_f32 else out.data_ptr(), out.data_ptr() if out_f32 else None,
            B, Z, Y, X, cin_p, cout, cout_p, int(relu), stream_of(x)))
    return out


class Conv3dFn(Function):
    """3x3x3 "same" convolution of channel-last bf16 activations."""

    @staticmethod
    def forward(ctx, x, w, b, relu: bool, cout_p: int):
        require_device(x, w, op="conv3d")
        if x.dtype != torch.bfloat16 or not x.is_contiguous():
            raise TypeError("conv3d expects contiguous channel-last bfloat16 activations")
        cout = w.shape[0]
        out = _fwd(x, pack_conv_weights(w, x.shape[-1]), b.detach().float().contiguous(),
                   cout, cout_p, relu)
        ctx.save_for_backward(x, w, out)
        ctx.relu = relu
        return out

    @staticmethod
    def backward(ctx, g):
        x, w, out = ctx.saved_tensors
        cout, cin = w.shape[:2]
        B, Z, Y, X, cin_p = x.shape
        cout_p = out.shape[-1]
        g = g.float() * (out > 0) if ctx.relu else g.float().clone()
        g[..., cout:] = 0
        gb = g.to(torch.bfloat16).contiguous()
        dx = _fwd(gb, pack_conv_weights_transposed(w, cout_p), None, cin, cin_p, False)
        ws_bytes = int(lib().tnl_conv3d_wgrad_workspace(B, Z, Y, X))
        ws = torch.empty(ws_bytes // 4, device=x.device, dtype=torch.float32)
        dw = torch.empty(cout, cin_p, 27, device=x.device, dtype=torch.float32)
        with torch.cuda.device(x.device):
            check(lib().tnl_conv3d_wgrad(x.data_ptr(), gb.data_ptr(), dw.data_ptr(),
                                         ws.data_ptr(), ws_bytes, B, Z, Y, X, cin_p, cout,
                                         cout_p, stream_of(x)))
        dw = dw[:, :cin].reshape(w.shape)
        db = gb[..., :cout].float().sum(dim=(0, 1, 2, 3))
        return dx, dw, db, None, None


def conv3d(x, w, b, relu: bool = False, cout_p: Optional[int] = None):
    return Conv3dFn.apply(x, w, b, relu, cout_p or _cpad(w.shape[0]))


class PrecondNet(nn.Module):
    """conv(F->16) ReLU conv(16->16) ReLU conv(16->16) ReLU conv(16->1);
    output o gives d = exp(o + log_scale)."""

    def __init__(self, in_features: int = 2, hidden: int = HIDDEN, layers: int = 4,
                 seed: int = 0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        chans = [in_features] + [hidden] * (layers - 1) + [1]
        self.weights = nn.ParameterList()
        self.biases = nn.ParameterList()
        for a, b in zip(chans[:-1], chans[1:]):
            std = math.sqrt(2.0 / (27 * a))
            self.weights.append(nn.Parameter(torch.randn(b, a, 3, 3, 3, generator=g) * std))
            self.biases.append(nn.Parameter(torch.zeros(b)))
        with torch.no_grad():
            self.weights[-1].zero_()  # output 0: the correction starts at 1
        self.log_scale = nn.Parameter(torch.zeros(()))

    def forward(self, feats: torch.Tensor) -> torch.Tensor:
        """feats fp32 [B, F, Z, Y, X] -> multiplicative correction fp32
        [B, Z, Y, X] (> 0; exactly 1 before training)."""
        x = to_channel_last(feats)
        n = len(self.weights)
        for i, (w, b) in enumerate(zip(self.weights, self.biases)):
            x = conv3d(x, w, b, relu=i < n - 1, cout_p=HIDDEN if i < n - 1 else 8)
        o = x[..., 0].float()
        return torch.exp(torch.clamp(o + self.log_scale, -30.0, 30.0))


def coil_features(maps: torch.Tensor) -> torch.Tensor:
    """[C, *im] complex coil maps -> [1, 2, *im] fp32: coil energy
    sum_c |S_c|^2 relative to its maximum, and its log."""
    e = (maps.abs() ** 2).sum(0)
    e = e / e.max()
    return torch.stack([e, torch.log(e + 1e-3) / math.log(1e3)], 0)[None].float()


def precond_inputs(A: "SenseNormal"):
    """(features [1, 2, *im] fp32, Jacobi diagonal J [*im] fp32) of a SENSE
    normal operator: J = 1 / (psf0 e + lam) with e = sum_c |S_c|^2 and psf0
    the diagonal of F^H F (constant over the image; measured at the centre
    voxel with a unit impulse); features q = psf0 e / (psf0 e + lam), the
    fraction of each voxel's diagonal that is signal, and log(J lam) /
    log(1e-4), both O(1).  Cached on the operator."""
    cached = getattr(A, "_precond_inputs", None)
    if cached is not None:
        return cached
    e = (A.maps.abs() ** 2).sum(0).float()
    c = tuple(n // 2 for n in A.im_size)
    delta = torch.zeros(A.im_size, dtype=torch.complex64, device=A.maps.device)
    delta[c] = 1.0
    with torch.no_grad():
        psf0 = float(A.normal_no_reg(delta)[c].real / e[c].clamp_min(1e-30))
    lam = A.lam if A.lam > 0 else 1e-6 * psf0 * float(e.max())
    diag = psf0 * e + lam
    J = 1.0 / diag
    q = psf0 * e / diag
    feats = torch.stack([q, torch.log(J * lam) / math.log(1e-4)], 0)[None].float()
    A._precond_inputs = (feats, J.float())
    return A._precond_inputs


def learned_diagonal(net: "PrecondNet", A: "SenseNormal") -> torch.Tensor:
    """d = J x net(features): the network's correction of Jacobi."""
    feats, J = precond_inputs(A)
    return J * net(feats)[0]


class SenseNormal:
    """x -> S^H F^H F S x + lam x for one image [*im] or a batch [R, *im]."""

    def __init__(self, locs: torch.Tensor, im_size: Sequence[int], maps: torch.Tensor,
                 oversamp: float = 2.0, width: float = 4.0, lam: float = 0.0):
        self.ndim = len(im_size)
        self.dim = tuple(range(-self.ndim, 0))
        self.im_size = tuple(int(s) for s in im_size)
        self.padded = tuple(int(s * oversamp) for s in self.im_size)
        self.width = float(width)
        self.beta = float(NUFFT.beta(width, oversamp))
        # no centring shifts: the image goes to the wrapped cells and the
        # locations are relabelled (see shifted_layout)
        self.index, shift = shifted_layout(self.im_size, self.padded, maps.device)
        pl = NUFFT.prep_locs(locs.reshape(-1, self.ndim), self.im_size, self.padded)
        pad = torch.tensor(self.padded, dtype=torch.float32, device=pl.device)
        self.locs = torch.remainder(pl - shift.to(pl.device), pad).contiguous()
        apod = NUFFT.apodize_weights(self.im_size, self.padded, oversamp, width, self.beta)
        scale = width**self.ndim * (math.prod(self.im_size) / math.prod(self.padded)) ** 0.5
        self.apod = apod.to(maps.device, torch.complex64)
        self.apod2 = (apod / scale**2).to(maps.device, torch.complex64)
        # without autograd: apodization (and 1/scale^2) with the orthonormal
        # FFT scalings folded in, applied by the wrapped pad / crop kernels
        rp = math.prod(self.padded) ** 0.5
        self.w_in = (apod / rp).to(maps.device, torch.float32).contiguous()
        self.w_out = (apod / (scale**2 * rp)).to(maps.device, torch.float32).contiguous()
        self.maps = maps.to(torch.complex64)
        self.lam = float(lam)
        self.kp = {"beta": self.beta}

    def normal_no_reg(self, x: torch.Tensor) -> torch.Tensor:
        lead = x.shape[: x.ndim - self.ndim]
        if not (torch.is_grad_enabled() and x.requires_grad):
            sx = self.maps * x.unsqueeze(-self.ndim - 1)  # [..., C, *im]
            g = pad_wrapped(sx, self.im_size, self.padded, weight=self.w_in)
            g = torch.fft.fftn(g, dim=self.dim, norm="backward")  # unscaled
            k = interpolate(g, self.locs, self.width, kernel_params=self.kp)
            g = interpolate_adjoint(k, self.locs, self.padded, self.width, kernel_params=self.kp)
            g = torch.fft.ifftn(g, dim=self.dim, norm="forward")  # unscaled
            g = crop_wrapped(g, self.im_size, self.padded, weight=self.w_out)
            return (self.maps.conj() * g).sum(-self.ndim - 1).reshape(*lead, *self.im_size)
        sx = self.maps * x.unsqueeze(-self.ndim - 1) * self.apod  # [..., C, *im]
        g = torch.zeros(*sx.shape[: sx.ndim - self.ndim], *self.padded, dtype=torch.complex64,
                        device=x.device)
        g[(Ellipsis,) + self.index] = sx
        g = torch.fft.fftn(g, dim=self.dim, norm="ortho")
        k = interpolate(g, self.locs, self.width, kernel_params=self.kp)
        g = interpolate_adjoint(k, self.locs, self.padded, self.width, kernel_params=self.kp)
        g = torch.fft.ifftn(g, dim=self.dim, norm="ortho")[(Ellipsis,) + self.index] * self.apod2
        out = (self.maps.conj() * g).sum(-self.ndim - 1)
        return out.reshape(*lead, *self.im_size)

    def __call__(self, x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        y = self.normal_no_reg(x)
        if self.lam:
            y = y + self.lam * x
        if out is not None:
            out.copy_(y)
            return out
        return y

    def normalize(self, iters: int = 20, seed: int = 0) -> float:
        """Rescale the coil maps so the largest eigenvalue of S^H F^H F S is
        1 (power iteration); lam is then relative to it.  Returns the old
        largest eigenvalue."""
        g = torch.Generator(device=self.maps.device).manual_seed(seed)
        v = torch.randn(self.im_size, generator=g, device=self.maps.device, dtype=torch.complex64)
        v = v / torch.linalg.vector_norm(v)
        lmax = 1.0
        for _ in range(iters):
            w = self.normal_no_reg(v)
            lmax = float(torch.linalg.vector_norm(w))
            v = w / lmax
        self.maps = self.maps / math.sqrt(lmax)
        return lmax


def pcg_unrolled(A, z: torch.Tensor, d: torch.Tensor, iters: int) -> torch.Tensor:
    """``iters`` preconditioned-CG iterations on A x = z from x = 0 with
    M = diag(d), as differentiable torch operations (batch of right-hand
    sides along the leading axis, per-sample scalars)."""
    nd = A.ndim
    red = tuple(range(z.ndim - nd, z.ndim))

    def dot(a, b):
        return (a.conj() * b).real.sum(dim=red, keepdim=True)

    x = torch.zeros_like(z)
    r = z
    w = d * r
    p = w
    rw = dot(r, w)
    for _ in range(iters):
        Ap = A(p)
        alpha = rw / dot(p, Ap).clamp_min(1e-30)
        x = x + alpha * p
        r = r - alpha * Ap
        w = d * r
        rw_new = dot(r, w)
        p = w + (rw_new / rw.clamp_min(1e-30)) * p
        rw = rw_new
    return x


def train_preconditioner(net: PrecondNet, data, steps: int = 300, lr: float = 3e-3,
                         unroll: int = 12, log=None, max_grad_norm: float = 1.0,
                         keep_best: bool = True):
    """data: list of (features, A, z [R, *im], x* [R, *im]) with x* = A^-1 z
    from plain CG (the features entry is unused: the diagonal is
    ``learned_diagonal(net, A)``).  Minimises the relative error of
    ``unroll`` preconditioned-CG iterations (Adam, gradients clipped to
    ``max_grad_norm``).  The loss through unrolled CG is rough, so with
    ``keep_best`` the parameters at the start of the pass over ``data``
    with the lowest mean loss are restored at the end.  Returns the
    per-step losses."""
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    losses = []
    nd = len(data)
    best, best_state, snap = float("inf"), None, None

    def close_pass():
        nonlocal best, best_state
        if snap is not None and len(losses) >= nd:
            pass_loss = sum(losses[-nd:]) / nd
            if pass_loss < best:
                best, best_state = pass_loss, snap

    for it in range(steps):
        _, A, z, x = data[it % nd]
        if keep_best and it % nd == 0:
            close_pass()
            snap = {k: v.detach().clone() for k, v in net.state_dict().items()}
        d = learned_diagonal(net, A)
        xk = pcg_unrolled(A, z, d, unroll)
        err = (xk - x).abs() ** 2
        red = tuple(range(1, x.ndim))
        loss = (err.sum(dim=red) / (x.abs() ** 2).sum(dim=red)).mean()
        opt.zero_grad()
        loss.backward()
        if max_grad_norm:
            torch.nn.utils.clip_grad_norm_(net.parameters(), max_grad_norm)
        opt.step()
        losses.append(float(loss.detach()))
        if log and (it % 50 == 0 or it == steps - 1):
            log(f"train step {it}: loss {losses[-1]:.4f}")
    if keep_best:
        if steps % nd == 0:
            close_pass()
        if best_state is not None:
            net.load_state_dict(best_state)
    return losses


def judge(A, y: torch.Tensor, M, iters: int = 10, check_every: int = 5):
    """Plain CG and CG preconditioned by ``M`` (r -> M r) from zero,
    compared at matched cost: ``iters`` preconditioned iterations against
    as many plain iterations as run in the same time (``iters``, or more
    when applying ``M`` makes a preconditioned iteration dearer; wall time
    of each run, synchronised).  Both are scored by the quantity CG
    minimises, the energy f(x) = x^H A x / 2 - Re(y^H x) = ||x - x*||_A^2 / 2
    - const (so a lower f is a smaller A-norm error, without knowing x*),
    computed from the same A x as the recomputed residual |y - A x| / |y|
    (reported too; CG's residual norm is not monotone on ill-conditioned
    systems, its energy is).  Returns (accept, cg_plain, cg_pre, info):
    the two ``DeviceCG`` runs, stopped where they were compared, so the
    winner continues its own Krylov sequence (no restart); ``accept`` is
    True when the preconditioned iterate is finite and its energy not above
    plain CG's."""
    ynorm = torch.linalg.vector_norm(y)
    sync = y.is_cuda

    def score(x):
        ax = A(x)
        f = 0.5 * float(torch.vdot(x.flatten(), ax.flatten()).real) - float(
            torch.vdot(y.flatten(), x.flatten()).real)
        return f, float(torch.linalg.vector_norm(y - ax) / ynorm)

    def timed(cg, n):
        if sync:
            torch.cuda.synchronize()
        t = time.perf_counter()
        cg.resume(tol=0.0, max_iters=cg.iters + n)
        if sync:
            torch.cuda.synchronize()
        return time.perf_counter() - t

    cg_m = DeviceCG(A, lam=0.0, M=M, check_every=check_every)
    cg_m.start(y)
    t_m = timed(cg_m, iters)
    cg_p = DeviceCG(A, lam=0.0, M=None, check_every=check_every)
    cg_p.start(y)
    t_p = timed(cg_p, iters)
    if t_p > 0 and t_m > 1.05 * t_p:
        # the preconditioner costs time of its own: plain CG runs on for
        # the time it leaves (at most 10x the budget)
        extra = min(int(round(iters * (t_m / t_p - 1.0))), 10 * iters)
        if extra > 0:
            t_p += timed(cg_p, extra)
    fp, rp = score(cg_p.x)
    fm, rm = score(cg_m.x)
    ok = math.isfinite(fm) and math.isfinite(rm) and fm <= fp
    info = {"accepted": ok, "iters": iters, "plain_iters": cg_p.iters,
            "precond_seconds": t_m, "plain_seconds": t_p, "plain_energy": fp,
            "precond_energy": fm, "plain_residual": rp, "precond_residual": rm}
    return ok, cg_p, cg_m, info


def solve(A: SenseNormal, y: torch.Tensor, d: Optional[torch.Tensor] = None,
          tol: float = 1e-6, max_iters: int = 1000, check_every: int = 10,
          check_factor: float = 100.0, judge_iters: int = 0):
    """(Preconditioned) CG until its relative residual |r| / |y| is below
    ``tol``.  Plain CG checks the result: with ``judge_iters`` > 0 the
    preconditioner must first match plain CG at matched cost (``judge``;
    else plain CG carries on from its own state);
    afterwards the residual is recomputed as |y - A x| / |y| with
    the operator itself, and if the preconditioned run did not converge,
    broke down (non-finite values) or its recomputed residual exceeds
    ``check_factor * tol``, plain CG continues from its iterate (or from
    zero).  Returns (x, info); info["iters"] counts every iteration run,
    the judge's included (the winner's judge iterations are the start of
    its own run)."""
    if y.is_cuda:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    # d: None (plain CG), an image-space diagonal, or a callable r -> M r
    M = None if d is None else (d if callable(d) else (lambda r: d * r))
    spent, verdict = 0, None
    if M is not None and judge_iters > 0:
        ok, cg_p, cg_m, verdict = judge(A, y, M, judge_iters, check_every=judge_iters)
        # the winner carries on from where the judge stopped it; the
        # loser's iterations count as spent
        cg, lost = (cg_m, cg_p) if ok else (cg_p, cg_m)
        if not ok:
            M = None
        spent = lost.iters
        cg.check_every = check_every
    else:
        cg = DeviceCG(A, lam=0.0, M=M, check_every=check_every)
        cg.start(y)
    x = cg.resume(tol=tol, max_iters=max_iters)
    iters = cg.iters + spent
    converged = cg.history[-1] <= tol
    ynorm = torch.linalg.vector_norm(y)
    true_res = float(torch.linalg.vector_norm(y - A(x)) / ynorm)
    fallback = False
    if M is not None and not (converged and math.isfinite(true_res)
                              and true_res <= check_factor * tol):
        fallback = True
        x0 = x if math.isfinite(true_res) else None
        cg2 = DeviceCG(A, lam=0.0, M=None, check_every=check_every)
        x = cg2.solve(y, x0=x0, tol=tol, max_iters=max_iters)
        iters += cg2.iters
        converged = cg2.history[-1] <= tol
        true_res = float(torch.linalg.vector_norm(y - A(x)) / ynorm)
    if y.is_cuda:
        torch.cuda.synchronize()
    return x, dict(seconds=time.perf_counter() - t0, iters=iters, converged=converged,
                   true_residual=true_res, fallback=fallback, history=list(cg.history),
                   judge=verdict)
