"""GPU: independent clouds evaluated as one recorded HIP graph (PtFMMBatch), and the per-tree
launch-shape autotuner.  The graph replays the same kernels on the same data as the eager
evaluations, so replays must equal them bit for bit; the graph is replayed several times with
eager work, in-place density changes and synchronizations between the replays, and every
replay is checked.  Retuned launch shapes change only the work split (the Hadamard run length
keeps every summation order; a different K split of a GEMM reorders its sums), so results
after tuning match to rounding."""
import numpy as np
import pytest

import pvfmm_amd as P

pytestmark = pytest.mark.gpu


def clouds(sizes, seed):
    rng = np.random.default_rng(seed)
    return [(rng.random((n, 3)), rng.standard_normal(n)) for n in sizes]


def test_batch_graph_replays_match_eager():
    import torch

    fmm = P.PtFMM("laplace_poten", 6, "mixed")
    cs = clouds([3000, 9000, 20000], 7)
    trees = [P.PtFMMTree(x, x, 64).setup(fmm) for x, _ in cs]
    dens = [torch.tensor(q, device="cuda:0") for _, q in cs]
    outs = [torch.empty_like(d) for d in dens]
    batch = P.PtFMMBatch(trees)
    for rep in range(4):
        print(f"replay {rep}", flush=True)
        batch.evaluate_device(dens, outs)
        ref = [t.evaluate_device(d.clone()) for t, d in zip(trees, dens)]
        for i, (o, r) in enumerate(zip(outs, ref)):
            assert torch.equal(o, r), f"replay {rep} cloud {i}: max diff {float((o - r).abs().max())}"
        # eager work on a tree of the batch, then new density values in the same buffers
        trees[0].evaluate_device(torch.ones_like(dens[0]))
        torch.cuda.synchronize()
        for d in dens:
            d.mul_(-0.5).add_(0.25)
    assert batch.records() == 1
    # other output buffers: the graph is recorded again
    outs2 = [torch.empty_like(d) for d in dens]
    batch.evaluate_device(dens, outs2)
    assert batch.records() == 2
    for t, d, o in zip(trees, dens, outs2):
        assert torch.equal(o, t.evaluate_device(d.clone()))
    # accuracy of a batched cloud against the direct sum
    x, _ = cs[1]
    ud = P.direct("laplace_poten", x, dens[1].cpu().numpy(), x)
    err = float(np.linalg.norm(outs2[1].cpu().numpy() - ud) / np.linalg.norm(ud))
    print(f"cloud 1 rel-L2 vs direct {err:.2e}")
    assert err < 1e-4


@pytest.mark.parametrize("prec,tol", [("mixed", 1e-6), ("fp64", 1e-12)])
def test_autotune_keeps_results(prec, tol):
    import torch

    rng = np.random.default_rng(3)
    x = rng.random((40000, 3))
    q = torch.tensor(rng.standard_normal(len(x)), device="cuda:0")
    t = P.PtFMMTree(x, x, 64).setup(P.PtFMM("laplace_grad", 6, prec))
    assert t.tuning() == {"hadamard_tiles_per_block": 16, "gemm_split_blocks": 1024, "hadamard_vec_targets_per_wave": 16}
    before = t.evaluate_device(q)
    chosen = t.autotune(q)
    print(f"{prec}: autotuned {chosen}")
    assert chosen["hadamard_tiles_per_block"] in (4, 8, 16, 32, 64)
    assert chosen["gemm_split_blocks"] in (0, 256, 512, 1024, 2048)
    assert t.tuning() == chosen
    after = t.evaluate_device(q)
    r = float(torch.linalg.norm(after - before) / torch.linalg.norm(before))
    assert r < tol, r
    # explicit extreme shapes: one tile per block run, no K split
    t.set_tuning(hadamard_tiles_per_block=1, gemm_split_blocks=0)
    r = float(torch.linalg.norm(t.evaluate_device(q) - before) / torch.linalg.norm(before))
    assert r < tol, r
    # runs longer than the FP32 Hadamard's 64-tile limit are clamped to it; the run length
    # does not change any sum, so the result is the same bit for bit as with 64
    t.set_tuning(hadamard_tiles_per_block=64, gemm_split_blocks=0)
    r64 = t.evaluate_device(q)
    t.set_tuning(hadamard_tiles_per_block=1000, gemm_split_blocks=0)
    assert torch.equal(t.evaluate_device(q), r64)


def test_batch_records_again_after_setup_and_orders_eager_calls():
    """A tree of the batch set up again (another PtFMM: a new plan with new device buffers)
    makes the next batch call record the graph again; replays and eager evaluations of the same
    trees, queued without synchronizing (sync=False), are ordered (they share the plans'
    scratch buffers) and each equals its own eager reference."""
    import torch

    cs = clouds([4000, 12000], 9)
    fmm6 = P.PtFMM("laplace_poten", 6, "mixed")
    trees = [P.PtFMMTree(x, x, 64).setup(fmm6) for x, _ in cs]
    dens = [torch.tensor(q, device="cuda:0") for _, q in cs]
    outs = [torch.empty_like(d) for d in dens]
    batch = P.PtFMMBatch(trees)
    batch.evaluate_device(dens, outs)
    assert batch.records() == 1
    trees[1].setup(P.PtFMM("laplace_poten", 8, "mixed"))
    batch.evaluate_device(dens, outs)
    assert batch.records() == 2
    ref = [t.evaluate_device(d.clone()) for t, d in zip(trees, dens)]
    for o, r in zip(outs, ref):
        assert torch.equal(o, r)
    # async: replay, eager evaluation of other densities on tree 0, replay again, all queued
    other = dens[0] * -2.0 + 1.0
    ref_other = trees[0].evaluate_device(other.clone())
    torch.cuda.synchronize()
    for rep in range(3):
        for o in outs:
            o.zero_()
        o_other = torch.zeros_like(other)
        torch.cuda.synchronize()
        batch.evaluate_device(dens, outs, sync=False)
        trees[0].evaluate_device(other, out=o_other, sync=False)
        batch.evaluate_device(dens, outs, sync=False)
        torch.cuda.ExternalStream(batch.stream_handle()).synchronize()
        for t in trees:
            torch.cuda.ExternalStream(t.stream_handle()).synchronize()
        assert torch.equal(o_other, ref_other), rep
        for o, r in zip(outs, ref):
            assert torch.equal(o, r), rep
    assert batch.records() == 2


@pytest.mark.parametrize("kernel", ["laplace_poten", "laplace_grad"])
def test_sixteen_vector_pass_and_its_tuning(kernel):
    """A plan set up for 8 or more vectors runs the V list of a pass through the 16-vector
    Hadamard (the vectors of one target in the MFMA rows).  Its products and their order are
    those of the one-vector kernel (absent partners add zero products), so every vector equals
    its own one-vector evaluation to rounding of the other phases; targets per wave 0 (one
    vector at a time through the tiled kernel) gives the same values; the nv > 1 autotuner picks
    between them and keeps the results."""
    import torch

    rng = np.random.default_rng(8)
    x = np.concatenate([rng.random((20000, 3)), 0.3 + 0.05 * rng.random((10000, 3))])
    n = len(x)
    q = torch.tensor(rng.standard_normal((20, n)), device="cuda:0")
    fmm = P.PtFMM(kernel, 6, "mixed")
    tb = P.PtFMMTree(x, x, 64).setup(fmm, n_vec_max=20)
    t1 = P.PtFMMTree(x, x, 64).setup(fmm, n_vec_max=1)
    ub = tb.evaluate_device(q)  # one pass: a group of 16 and a group of 4
    worst = 0.0
    for v in range(20):
        u1 = t1.evaluate_device(q[v].contiguous())
        worst = max(worst, float(torch.linalg.norm(ub[v] - u1) / torch.linalg.norm(u1)))
    print(f"{kernel}: 20 vectors in one pass vs one at a time, worst rel {worst:.2e}")
    assert worst < 1e-6
    # targets per wave 0: the same pass one vector at a time through the tiled kernel
    tb.set_tuning(hadamard_vec_targets_per_wave=0)
    u0 = tb.evaluate_device(q)
    assert float(torch.linalg.norm(u0 - ub) / torch.linalg.norm(ub)) < 1e-6
    chosen = tb.autotune(q[:16].contiguous())
    print(f"autotuned for 16 vectors: {chosen}")
    assert chosen["hadamard_vec_targets_per_wave"] in (0, 4, 8, 16, 32)
    ub2 = tb.evaluate_device(q)
    assert float(torch.linalg.norm(ub2 - ub) / torch.linalg.norm(ub)) < 1e-6
