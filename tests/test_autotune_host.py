"""Host logic of the particle-step autotuner, no GPU: a fake context and
state whose kernel times follow the parameters set (a cost model), so that
each sweep's choice is known: the push tile and register budget, the deposit
variant (a pass that cannot read a step's particles through the migration's
source indices wins only if whole steps are faster with it), the node batch,
the report fields and the cache."""
import json

import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import autotune as A

PUSH = {1024: 5.0, 2048: 4.5, 4096: 4.0, 8192: 4.2}
NODE = {4: 4.4, 8: 4.1, 16: 4.0}
COMPACT = 0.45       # a step's migration that leaves the records for the next push
SCATTER = 2.3        # the record scatter of a step whose cell pass cannot read through the indices
RECORD_MOVE = 0.3    # the next push's share of the record move


class FakeCtx:
    def __init__(self):
        self.p = {}
        self.on = False
        self.reset_timing()

    def set_param(self, name, value):
        self.p[name] = int(value)

    def set_timing(self, on):
        self.on = bool(on)

    def reset_timing(self):
        self.ms = np.zeros(L.K_COUNT)
        self.n = np.zeros(L.K_COUNT, np.int64)

    def timing(self):
        return self.ms.copy(), self.n.copy()

    def add(self, k, ms):
        assert self.on, "kernel times are read only with timing on"
        self.ms[k] += ms
        self.n[k] += 1


class FakeState:
    """Kernel times from the context's parameters; deposit[(kernel, chunks)]
    is the cell pass."""

    def __init__(self, ctx, deposit):
        self.ctx = ctx
        self.deposit = deposit

    def _push(self, move=False):
        p = self.ctx.p
        self.ctx.add(L.K_PUSH, PUSH[p["tile"]] - (0.1 if p["push_waves"] == 4 else 0.0)
                     + (RECORD_MOVE if move else 0.0))

    def _deposit(self):
        p = self.ctx.p
        self.ctx.add(L.K_DEPOSIT_CELL, self.deposit[(p["deposit_kernel"], p["deposit_chunks"])])
        self.ctx.add(L.K_DEPOSIT_NODE, NODE[p.get("node_batch", 8)])

    def MoveParticles(self, compact=True, **kw):
        self._push()

    def UpdateJMassMatrix(self, dt, copy_out=True):
        self._deposit()

    def run_steps(self, nsteps, dt, use_graph=True, **kw):
        fused = self.ctx.p["deposit_kernel"] in (2, 3)
        for _ in range(nsteps):
            self._push(move=fused)
            self.ctx.add(L.K_COMPACT, COMPACT + (0.0 if fused else SCATTER))
            self._deposit()


def _deposit_times(best=None):
    t = {v: 9.0 + 0.5 * i for i, v in enumerate(A.DEPOSIT_VARIANTS)}
    t.update(best or {})
    return t


def _tune(tmp_path, deposit):
    ctx = FakeCtx()
    st = FakeState(ctx, deposit)
    cache = str(tmp_path / "tune.json")
    best, cached = A.tune_pic(ctx, st, {}, 0.1, "sig", cache=cache)
    return ctx, st, cache, best, cached


def test_fusable_deposit_winner_needs_no_step_comparison(tmp_path):
    ctx, st, cache, best, cached = _tune(tmp_path, _deposit_times())
    # the deposit variants' own times rise along the list: (3, 8) wins
    assert not cached
    assert (best["tile"], best["push_waves"]) == (4096, 4)
    assert (best["deposit_kernel"], best["deposit_chunks"]) == A.DEPOSIT_VARIANTS[0] == (3, 8)
    assert best["node_batch"] == 16 and best["step_ms"] == {}
    assert set(best["deposit_ms"]) == {f"k{k}c{c}" for k, c in A.DEPOSIT_VARIANTS}
    assert best["push_ms"] == {str(t): PUSH[t] for t in A.PUSH_TILES}
    assert ctx.p["tile"] == 4096 and ctx.p["push_waves"] == 4 and ctx.p["node_batch"] == 16
    assert (ctx.p["deposit_kernel"], ctx.p["deposit_chunks"]) == (3, 8) and not ctx.on


@pytest.mark.parametrize("margin,winner", [(1.0, (2, 4)), (3.0, (1, 4))])
def test_other_deposit_winner_must_win_whole_steps(tmp_path, margin, winner):
    """(1, 4) is the fastest deposit on its own, by `margin` over (2, 4), the
    fastest pass that reads through the indices; its steps pay the record
    scatter (SCATTER - RECORD_MOVE = 2.0 ms), so it keeps the choice only
    when its margin exceeds that."""
    dep = _deposit_times({(1, 4): 5.0, (2, 4): 5.0 + margin})
    ctx, st, cache, best, cached = _tune(tmp_path, dep)
    assert (best["deposit_kernel"], best["deposit_chunks"]) == winner
    assert (ctx.p["deposit_kernel"], ctx.p["deposit_chunks"]) == winner
    steps = best["step_ms"]
    assert set(steps) == {"k1c4", "k2c4"}
    # steps are weighed before the node batch is tuned (the context's 8)
    push = PUSH[4096] - 0.1
    assert steps["k1c4"] == pytest.approx(push + COMPACT + SCATTER + 5.0 + NODE[8])
    assert steps["k2c4"] == pytest.approx(push + RECORD_MOVE + COMPACT + 5.0 + margin + NODE[8])


def test_cache_applies_the_stored_choice(tmp_path):
    dep = _deposit_times({(1, 4): 5.0, (2, 4): 6.0})
    _, _, cache, best, _ = _tune(tmp_path, dep)
    with open(cache) as f:
        assert json.load(f)["sig"] == best
    ctx = FakeCtx()
    again, cached = A.tune_pic(ctx, FakeState(ctx, dep), {}, 0.1, "sig", cache=cache)
    assert cached and again == best
    assert ctx.p == {"tile": best["tile"], "push_waves": best["push_waves"],
                     "deposit_kernel": best["deposit_kernel"], "deposit_chunks": best["deposit_chunks"],
                     "node_batch": best["node_batch"]}
