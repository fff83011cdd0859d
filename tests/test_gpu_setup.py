"""The set-up on the GPU (Ruiz and Pock-Chambolle passes, the transpose, the power iteration's SpMVs) against the
host's restatement of the reference's scale_lp + estimate_norm (src/scaling.cpp:36-155), which test_host.py checks
bit for bit against the oracle: every array and scalar of the scaled problem, and the norm estimate, must be the
same bits -- so a solve set up on the GPU follows the host-set-up solve iteration for iteration."""
import numpy as np
import pytest

import ganit_v2_amd as g
from conftest import instance_path

pytestmark = pytest.mark.gpu

ARRAYS = ["c", "l", "u", "lo", "hi", "R", "C", "Rres", "Cres"]
SCALARS = ["bscale", "cscale", "normQ", "obj_const", "bnorm", "cnorm", "normA", "nnz_at", "maximize"]


def _problems():
    return {
        "afiro": g.Problem.read_mps(instance_path("afiro")),
        "25fv47": g.Problem.read_mps(instance_path("25fv47")),
        "greenbea": g.Problem.read_mps(instance_path("greenbea")),
        "port_small": g.Problem.read_mps(instance_path("port_small")),
        "random_lp": g.Problem.generate("random_lp", seed=1, m=1500, n=4000, m_eq=700),
        "blending_small": g.Problem.generate("blending", seed=2, crudes=12, products=5, specs=3, periods=40),
    }


def _assert_same(h, d, what):
    for k in SCALARS:
        assert h[k] == d[k], (what, k, h[k], d[k])
    for k in ARRAYS:
        assert np.array_equal(h[k], d[k]), (what, k)
    for k in ("A", "AT", "Q"):
        for f in ("ptr", "idx", "val"):
            assert np.array_equal(getattr(h[k], f), getattr(d[k], f)), (what, k, f)


@pytest.mark.parametrize("name", list(_problems()))
def test_gpu_setup_is_bitwise_the_host_setup(name):
    p = _problems()[name]
    _assert_same(p.scaled(), p.scaled(on_gpu=True), name)


@pytest.mark.parametrize("ruiz,pc", [(0, False), (3, False), (10, True)])
def test_gpu_setup_pass_counts(ruiz, pc):
    p = g.Problem.generate("random_lp", seed=4, m=600, n=1500, m_eq=300)
    _assert_same(p.scaled(ruiz, pc), p.scaled(ruiz, pc, on_gpu=True), (ruiz, pc))


def test_gpu_setup_of_the_blending_lp_and_its_solve():
    """The 2.9M-nonzero blending LP: the same scaled problem and norm, and solves that agree bit for bit
    (iterations, restarts, objectives) whichever side set them up."""
    p = g.Problem.generate("blending", seed=0)
    _assert_same(p.scaled(), p.scaled(on_gpu=True), "blending")
    rs = {}
    for where in ("host", "gpu"):
        s = p.solver(precision="mixed", tol=1e-4, stall_checks=60, eval_fp64=True, setup=where)
        s.upload()
        rs[where] = s.run()
        print(f"setup {where}: {rs[where]['setup_s']:.4f} s; solve {rs[where]['status']} it {rs[where]['iters']} "
              f"obj {rs[where]['pobj']!r}", flush=True)
    h, d = rs["host"], rs["gpu"]
    assert h["status"] == d["status"] == "OPTIMAL"
    assert h["iters"] == d["iters"] and h["restarts"] == d["restarts"]
    assert h["pobj"] == d["pobj"] and h["dobj"] == d["dobj"]


def test_gpu_setup_of_an_empty_matrix():
    p = g.Problem.from_triplets(0, 3, [], [], [], c=[1.0, -1.0, 0.0], l=[0.0] * 3, u=[1.0] * 3, lo=[], hi=[])
    _assert_same(p.scaled(), p.scaled(on_gpu=True), "empty")


def test_automatic_setup_placement():
    """Auto places the set-up on the GPU from a million nonzeros (rows and columns of at most 4096 entries), on
    the host below that and for matrices with longer rows (the 100,000-wide factor rows of the portfolio QP)."""
    assert g.Problem.generate("blending", seed=0).solver().setup_on_device()
    assert not g.Problem.read_mps(instance_path("afiro")).solver().setup_on_device()
    assert not g.Problem.generate("portfolio", seed=0).solver().setup_on_device()
    assert g.Problem.read_mps(instance_path("afiro")).solver(setup="gpu").setup_on_device()
    assert not g.Problem.generate("blending", seed=0).solver(setup="host").setup_on_device()
