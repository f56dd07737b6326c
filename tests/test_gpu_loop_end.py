"""The end of a fused leapfrog loop with tensors (128 <= N <= 256, SIGW).

A fused step's tensor launch reads the background record that the step's march
launch writes.  Once a > af the march launches return at once; the record they
pass on must still say the loop has ended, or the tensor launches queued after
the end would step h and hdot again with an older record's scalars.  So a loop
that ends part way through a call of many steps (one chunk, many launches left)
gives the bits of the same loop taken one step per call, with calls after its
end changing nothing.  Before the fix the long call took an extra tensor step
on every other launch after the end, and whole runs of the same inputs gave
Omega_GW that depended on how the steps were split into calls.
"""
from __future__ import annotations

import numpy as np
import pytest

import ie_engine as E

pytestmark = pytest.mark.gpu

HILL = dict(V0=3.338e-13, ns=0.97, dt=0.001)


def new_run(N, precision, af):
    cfg = E.make_config(N, precision=precision, sigw=1, af=af, **HILL)
    run = E.Run(cfg)
    run.initialize(71, 0.0935, 0.000796)
    run.evolve_fields(0.5 * cfg.dt)
    return run


def snapshot(run):
    bg = run.background()
    return ([run.field(E.IE_FIELD_F), run.field(E.IE_FIELD_FD)] + [run.tensor(c) for c in range(12)],
            (bg["steps"], bg["a"], bg["ad"], bg["done"]))


def assert_same(a, b):
    (xa, ba), (xb, bb) = a, b
    assert ba == bb
    for u, v in zip(xa, xb):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("N,precision", [(128, E.IE_SPLIT), (256, E.IE_SPLIT), (128, E.IE_FP64)])
def test_loop_end_inside_a_call_leaves_tensors_alone(N, precision):
    af = 1.03
    print(f"loop end N={N} precision={precision}", flush=True)
    one = new_run(N, precision, af)
    one.set_shapes(0, 0, 0)
    bg = one.leapfrog_steps(400)
    assert bg["done"] == 1 and 10 < bg["steps"] < 200   # ends with > 200 launches left
    many = new_run(N, precision, af)
    many.set_shapes(0, 0, 0)
    while not many.leapfrog_steps(1)["done"]:
        pass
    many.leapfrog_steps(7)
    many.leapfrog_steps(1)
    assert_same(snapshot(one), snapshot(many))
    one.close()
    many.close()


def test_batched_loop_end_leaves_tensors_alone():
    af = 1.03
    print("batched loop end N=128 split", flush=True)
    one, many = new_run(128, E.IE_SPLIT, af), new_run(128, E.IE_SPLIT, af)
    E.Batch([one]).steps(400, graph_steps=0)
    b = E.Batch([many])
    for _ in range(400):
        b.steps(1, graph_steps=0)
        if many.background()["done"]:
            break
    b.steps(9, graph_steps=0)
    b.close()
    assert_same(snapshot(one), snapshot(many))
    assert one.background()["done"] == 1
    one.close()
    many.close()
