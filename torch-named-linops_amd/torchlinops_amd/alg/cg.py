"""Conjugate gradients.

``conjugate_gradients`` keeps the reference's interface and semantics
(src/torchlinops/alg/pcg.py:13-176): plain CG from x0 (default 0), at most
``max_num_iters`` steps, stopping when both |loss change| < ltol and
||A x - y|| < gtol, where tracking costs one extra A(x) and two host syncs
per iteration unless ``disable_tracking``.

``DeviceCG`` is the device-resident solver for (A + lam I) x = y used by
the reconstruction path: all scalars stay on the GPU, the relative
residual ||r|| / ||y|| is checked on the host only every ``check_every``
iterations, and a block of iterations can be captured once in a HIP graph
and replayed.  An optional preconditioner M (callable, approximating
(A + lam I)^-1) turns it into preconditioned CG.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import torch

__all__ = ["conjugate_gradients", "CGRun", "inner", "DeviceCG"]


def inner(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """Complex inner product sum(conj(x) * y)."""
    return torch.sum(x.conj() * y)


@dataclass
class CGRun:
    ltol: float
    gtol: float
    A: Callable
    y: torch.Tensor
    x_out: Optional[torch.Tensor] = None
    prev_loss: Optional[float] = None
    loss: float = float("inf")
    gnorm: Optional[float] = None
    disable: bool = False

    def update(self, x):
        if self.disable:
            self.x_out = x
            return
        self.prev_loss = self.loss
        Ax = self.A(x)
        xy = inner(x, self.y)
        self.loss = (inner(x, Ax) - xy - xy.conj()).real.item()
        self.x_out = x
        self.gnorm = torch.linalg.vector_norm(Ax - self.y).item()

    def is_converged(self) -> bool:
        if self.disable:
            return False
        return abs(self.loss - self.prev_loss) < self.ltol and self.gnorm < self.gtol

    def set_postfix(self, pbar):
        if not self.disable:
            pbar.set_postfix({"ldiff": abs(self.loss - self.prev_loss), "gnorm": self.gnorm})


def conjugate_gradients(A: Callable, y: torch.Tensor, x0: Optional[torch.Tensor] = None,
                        max_num_iters: int = 20, gtol: float = 1e-3, ltol: float = 1e-5,
                        disable_tracking: bool = False, tqdm_kwargs: Optional[dict] = None):
    """Solve A x = y (A Hermitian positive semidefinite)."""
    from tqdm import tqdm

    x = torch.zeros_like(y) if x0 is None else x0.clone()
    kw = dict(desc="CG", leave=False)
    kw.update(tqdm_kwargs or {})
    run = CGRun(ltol, gtol, A, y, disable=disable_tracking)
    run.update(x)
    r = y - A(x)
    p = r.clone()
    rs = inner(r, r).real
    with tqdm(range(max_num_iters), **kw) as pbar:
        for _ in pbar:
            Ap = A(p)
            alpha = rs / inner(p, Ap)
            x = x + alpha * p
            r = r - alpha * Ap
            rs_old = rs.clone()
            rs = inner(r, r).real
            run.update(x)
            if run.is_converged():
                break
            run.set_postfix(pbar)
            p = (rs / rs_old) * p + r
    return run.x_out


@dataclass
class DeviceCG:
    """Preconditioned CG for (A + lam I) x = y with device-side scalars.

    A: callable x -> A x (e.g. a linop's .N); M: optional callable r -> z
    applying the preconditioner.  ``solve`` runs until the relative
    residual drops below ``tol`` or ``max_iters``; ``history`` collects the
    relative residuals seen at each host check.
    """

    A: Callable
    lam: float = 0.0
    M: Optional[Callable] = None
    check_every: int = 5
    use_graph: bool = False
    history: list = field(default_factory=list)

    def _op(self, x):
        y = self.A(x)
        return y + self.lam * x if self.lam else y

    def _pre(self, r):
        return self.M(r) if self.M is not None else r

    def _init(self, y, x0):
        self.y = y
        self.x = torch.zeros_like(y) if x0 is None else x0.clone()
        self.r = y - self._op(self.x) if x0 is not None else y.clone()
        self.z = self._pre(self.r)
        self.p = self.z.clone()
        self.rz = inner(self.r, self.z).real
        self.ynorm = torch.linalg.vector_norm(y)

    def _iteration(self):
        Ap = self._op(self.p)
        alpha = self.rz / inner(self.p, Ap).real
        self.x.add_(alpha * self.p)
        self.r.sub_(alpha * Ap)
        z = self._pre(self.r)
        rz_new = inner(self.r, z).real
        beta = rz_new / self.rz
        self.p.mul_(beta).add_(z)
        self.rz.copy_(rz_new)

    def rel_residual(self) -> float:
        return float(torch.linalg.vector_norm(self.r) / self.ynorm)

    def solve(self, y: torch.Tensor, x0: Optional[torch.Tensor] = None, tol: float = 1e-6,
              max_iters: int = 200) -> torch.Tensor:
        self.start(y, x0)
        return self.resume(tol, max_iters)

    def start(self, y: torch.Tensor, x0: Optional[torch.Tensor] = None) -> None:
        """Set up the iteration for A x = y from x0 (zero when None)."""
        self._init(y, x0)
        self.history = [self.rel_residual()]
        self.iters = 0
        self._graph, self._graph_n = None, 0

    def resume(self, tol: float = 1e-6, max_iters: int = 200) -> torch.Tensor:
        """Continue the iteration from its current state (the same Krylov
        sequence as one uninterrupted run) until the relative residual is
        below ``tol`` or ``max_iters`` iterations in total have run.  The
        residual is checked at multiples of ``check_every`` iterations of
        the whole run (a run stopped between two checks catches up to the
        next multiple first)."""
        while self.iters < max_iters and self.history[-1] > tol:
            n = self.check_every - self.iters % self.check_every
            n = min(n, max_iters - self.iters)
            if self.use_graph and self.y.is_cuda and n == self.check_every:
                if self._graph is None or self._graph_n != n:
                    # runs this block eagerly, then records it
                    self._graph, self._graph_n = self._capture(n), n
                else:
                    self._graph.replay()
            else:
                for _ in range(n):
                    self._iteration()
            self.iters += n
            self.history.append(self.rel_residual())
        return self.x

    def _capture(self, n):
        """Run n iterations eagerly on a side stream (so plans and
        allocations exist), then record n iterations into a graph without
        executing them.  The plans the recorded kernels read are kept in
        ``self.graph_plans`` for as long as the graph is, so a plan-cache
        eviction cannot free a workspace the graph points to."""
        from ..functional._plan import record_plans

        self.graph_plans = []
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with record_plans(self.graph_plans):
            with torch.cuda.stream(s):
                for _ in range(n):
                    self._iteration()
            torch.cuda.current_stream().wait_stream(s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                for _ in range(n):
                    self._iteration()
        return g
