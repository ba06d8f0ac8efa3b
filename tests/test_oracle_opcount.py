"""The chemistry operation counts (oracle/opcount_table.py, the oracle's
chemistry templates run on a counting scalar) that the benchmark uses as the
algorithmic FP64 work of the chemistry kernels: the counter itself pinned by
the closed-form operation count of the pivoted solve, the Newton count's
iterations equal to the oracle's own iteration count, and the committed
table up to date."""
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def h2():
    import flame3d_amd.mech as mm
    return mm.load("h2_lidryer")


@pytest.mark.parametrize("n", [2, 3, 5, 10])
def test_gauss_opcount_closed_form(h2, n):
    """Full-row elimination (reactions.jl:463-503): for each pivot k and row
    i > k one division and n + 1 multiply-adds; back substitution n divisions
    and n(n-1)/2 multiply-adds."""
    from oracle import Oracle
    orc = Oracle(1, 1, 1, h2)
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + n * np.eye(n)
    c = orc.gauss_opcount(A, rng.normal(size=n))
    pairs = n * (n - 1) // 2
    assert c["div"] == pairs + n
    assert c["mul"] == (n + 1) * pairs + pairs
    assert c["add"] == (n + 1) * pairs + pairs
    assert c["trans"] == 0


def test_newton_opcount_iterations_match_oracle(h2):
    """The counted Newton solve takes the iterations the oracle's
    or_chem_newton reports for the same reactor (single step, no restart)."""
    import flame3d_amd.cases as cases
    from oracle import Oracle
    Q, r = cases.chem_cells((4, 4, 2), h2, seed=5, dev="cpu")
    orc = Oracle(4, 4, 2, h2)
    Qn = np.array(Q.numpy().T, order="F")
    rn = np.array(r.numpy().T, order="F")
    U = orc.prim2c(Qn)
    iters = orc.chem_newton(U, Qn.copy(order="F"), rn.copy(order="F"), 1e-7)
    o1 = Oracle(1, 1, 1, h2)
    for c in range(32):
        a, b, k = c % 4, (c // 4) % 4, c // 16
        cnt = o1.newton_opcount(r[:, k + 4, b + 4, a + 4].numpy().copy(), float(Q[5, k + 4, b + 4, a + 4]),
                                1e-7)
        it = int(iters[a, b, k])
        assert it > 0 and cnt["iterations"] == it, (c, cnt, it)
        # one iteration: two rate evaluations (one with the Jacobian) and
        # the 10x10 pivoted solve at least
        per = (cnt["add"] + cnt["mul"] + cnt["div"]) / it
        assert 3000 < per < 20000 and cnt["trans"] / it >= 2 * 21


def test_committed_table_is_current():
    import oracle.opcount_table as ot
    res = ot.compute()
    old = json.load(open(os.path.join(ROOT, "profiles", "chem_opcount.json")))
    assert old == json.loads(json.dumps(res))
