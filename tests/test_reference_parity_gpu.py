"""GPU results against the reference's own CPU kernels on the same inputs.

oracle/_ref holds Omega3D's kernel_0v_0bg and core_func (Kernels.h, CoreFunc.h)
compiled in place, driven by the target loop of points_affect_points
(Influence.h:405-440) with S = float and A = double: the reference's CPU build
without Vc.  The bar is 1e-5 relative (normwise max over targets), on
  - the single 8,192-particle ring (RM core), every target, plain FP32 and
    compensated, self pair included as the reference does, and self-excluded plus
    the closed-form self term K(0) skew(alpha);
  - the two colliding rings of 1,048,576 particles (EXP core), 1,024 spread
    targets of the full-collection call, by the direct and the factored method.
Skipped where oracle/_ref has not been built (it needs the reference sources).
"""
import numpy as np
import pytest

import oracle
from omega3d_amd import scenes
from omega3d_amd.metrics import rel_errors, self_term_gradient

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(oracle.ref_lib("rm") is None or oracle.ref_lib("exp") is None,
                                 reason="oracle/_ref (the reference's own kernels) is not built")]

BAR = 1e-5


@pytest.fixture(scope="module")
def infl():
    from omega3d_amd import _lib, influence
    assert _lib.lib.o3d_init(-1) == 0, _lib.last_error()
    return influence


def test_single_ring_every_target(infl):
    p = scenes.ring(8192, sigma=0.05)
    rv, rg = oracle.ref_p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "f32d")
    for prec in ("fp32", "comp"):
        v, g = infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, prec)
        ev, eg = rel_errors(v, rv), rel_errors(g, rg)
        print(f"single ring {prec}: vel {ev} grad {eg}", flush=True)
        assert ev[0] <= BAR and eg[0] <= BAR, prec
    # the workload as configured (self pair excluded) plus the self term is the reference
    v, g = infl.p2p("rm", p.pos, p.str, p.rad, p.pos, p.rad, True, "comp", self_term=1)
    g = g.astype(np.float64) + self_term_gradient("rm", p.str, p.rad)
    assert rel_errors(v, rv)[0] <= BAR and rel_errors(g, rg)[0] <= BAR


@pytest.fixture(scope="module")
def rings_ref():
    c = scenes.colliding_rings()
    sel = np.linspace(0, c.n - 1, 1024).astype(np.int64)
    rv, rg = oracle.ref_p2p("exp", c.pos, c.str, c.rad, c.pos[:, sel], c.rad[sel], True, "f32d")
    return c, sel, rv, rg


@pytest.mark.parametrize("method", ["direct", "factored"])
def test_two_rings_1m_sampled_targets(infl, rings_ref, method):
    c, sel, rv, rg = rings_ref
    print(f"two rings N={c.n} EXP compensated, method {method}", flush=True)
    v, g = infl.p2p("exp", c.pos, c.str, c.rad, c.pos, c.rad, True, "comp", method=method)
    ev, eg = rel_errors(v[:, sel], rv), rel_errors(g[:, sel], rg)
    print(f"  vs reference on {len(sel)} targets: vel {ev} grad {eg}", flush=True)
    assert ev[0] <= BAR and eg[0] <= BAR
