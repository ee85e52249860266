"""The plain-CG judge of the preconditioned solver, on the CPU.

``alg.learned.judge`` compares a preconditioned CG run with plain CG at
matched cost and ``solve`` lets the winner carry on its own Krylov sequence.
Checked on small dense SPD systems (no GPU needed):
  * ``DeviceCG.resume`` continues the same iteration: 7 + 13 iterations give
    the iterate of 20 uninterrupted ones, bit for bit;
  * an exact preconditioner (A^-1) is accepted and the solve converges in a
    few iterations, the judge's plain run counted;
  * a preconditioner that changes nothing but costs time of its own (the
    identity behind a sleep) is rejected, because at matched time plain CG
    runs more iterations; the solve then carries on with plain CG;
  * a rank-one preconditioner (search directions along one vector) is
    rejected at matched iteration counts.
"""

import time

import numpy as np
import torch

from torchlinops_amd.alg.cg import DeviceCG
from torchlinops_amd.alg import learned as Lm


def _spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1.0, cond, n)
    a = torch.from_numpy((q * lam) @ q.T)
    y = torch.from_numpy(rng.standard_normal(n))
    return a, y


def test_resume_continues_the_same_iteration():
    a, y = _spd(60, 1e3, 1)

    def op(x):
        return a @ x

    one = DeviceCG(op, check_every=1)
    x20 = one.solve(y, tol=0.0, max_iters=20).clone()
    two = DeviceCG(op, check_every=1)
    two.start(y)
    two.resume(tol=0.0, max_iters=7)
    two.check_every = 5
    x = two.resume(tol=0.0, max_iters=20)
    assert two.iters == 20
    assert torch.equal(x, x20)


def test_exact_preconditioner_is_accepted():
    a, y = _spd(80, 1e4, 2)
    ainv = torch.linalg.inv(a)
    _, plain = Lm.solve(lambda x: a @ x, y, None, tol=1e-10, max_iters=2000, check_every=1)
    _, info = Lm.solve(lambda x: a @ x, y, lambda r: ainv @ r, tol=1e-10, max_iters=2000,
                       check_every=1, judge_iters=3)
    j = info["judge"]
    assert j["accepted"] and j["precond_energy"] <= j["plain_energy"]
    assert info["converged"] and not info["fallback"]
    # 3 judge iterations of each: the preconditioned run is already at the
    # solution, so it stops at once; plain CG's 3 (or more) count as spent
    assert j["plain_iters"] <= info["iters"] < plain["iters"]


def test_costly_preconditioner_without_gain_is_rejected():
    a, y = _spd(80, 1e3, 3)

    def slow_identity(r):
        time.sleep(0.005)
        return r.clone()

    x, info = Lm.solve(lambda x: a @ x, y, slow_identity, tol=1e-8, max_iters=2000,
                       check_every=1, judge_iters=4)
    j = info["judge"]
    # identical iterates at matched counts; at matched time plain CG has run
    # further, so its energy is lower
    assert j["plain_iters"] > j["iters"]
    assert not j["accepted"] and j["plain_energy"] < j["precond_energy"]
    assert info["converged"] and not info["fallback"]
    assert float(torch.linalg.vector_norm(y - a @ x) / torch.linalg.vector_norm(y)) < 1e-6


def test_rank_one_preconditioner_is_rejected():
    a, y = _spd(50, 1e2, 4)
    v = torch.from_numpy(np.random.default_rng(5).standard_normal(50))
    v = v / torch.linalg.vector_norm(v)
    x, info = Lm.solve(lambda x: a @ x, y, lambda r: v * torch.dot(v, r), tol=1e-8,
                       max_iters=2000, check_every=1, judge_iters=6)
    j = info["judge"]
    assert not j["accepted"] and j["precond_energy"] > j["plain_energy"]
    assert info["converged"]
