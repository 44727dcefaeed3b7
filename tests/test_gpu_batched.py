"""Batched form vs plain per-operator launches on the same pipelines (GPU): group keys and counts
bit-exact, float64 sums within 1e-12 relative."""
import copy

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _plain(group, plan):
    from tqp_amd import executor
    tg = dict(group)
    executor.execute(copy.deepcopy(plan), tg)
    return tg


@pytest.mark.parametrize("records,fast", [(True, True), (False, True), (True, False),
                                          (False, False)])
def test_batched_matches_plain(device, records, fast):
    """records=True: tables beyond LDS aggregate through records and slice reduces; False: with
    global atomics.  fast=True: conjunctive predicates over whole tiles; False: the per-row
    interpreter."""
    from tqp_amd import batched, workloads
    probe, build, consts = workloads.config3_sizes(n_pipes=24, lo=1000, hi=300_000)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], device)
             for i in range(len(probe))]
    specs = [batched.match_plan(plan, group) for group, plan in pipes]
    assert all(s is not None for s in specs)
    eng = batched.BatchedEngine(specs, tile_rows=2048, records=records, fast=fast)
    nrec = sum(int(d.records) for d in eng._desc_host[:eng.npipes])
    print("records pipelines:", nrec, "of", eng.npipes, "fast predicates:",
          eng.fast_predicates, flush=True)
    assert (nrec > 0) == records
    assert (eng.fast_predicates == 2 * eng.npipes) == fast
    for rep in range(2):                     # a replay must give the same results
        eng.run()
        out, rerun = eng.results()
        assert rerun == []
        for i, (group, plan) in enumerate(pipes):
            ref = _plain(group, plan)
            got = out[i]
            print("pipeline", i, "groups", got[101].shape[0], flush=True)
            assert torch.equal(got[101].cpu(), ref[101].tensor.cpu()), f"keys of pipeline {i}"
            assert torch.equal(got[302].cpu(), ref[302].tensor.cpu()), f"counts of pipeline {i}"
            np.testing.assert_allclose(got[301].cpu().numpy(), ref[301].tensor.cpu().numpy(),
                                       rtol=1e-12, atol=0, err_msg=f"sums of pipeline {i}")


@pytest.mark.parametrize("tmax,rlds", [(16 * 1024, 16 * 1024), (48 * 1024, 32 * 1024)])
def test_batched_lds_capacities(device, tmax, rlds):
    """The LDS capacities the autotuner chooses among (the whole-table limit, the records slice
    table) change where pipelines accumulate, not their results."""
    from tqp_amd import batched, workloads
    probe, build, consts = workloads.config3_sizes(n_pipes=12, lo=1000, hi=200_000)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], device)
             for i in range(len(probe))]
    specs = [batched.match_plan(plan, group) for group, plan in pipes]
    eng = batched.BatchedEngine(specs, tile_rows=2048, records=True, lds_table_max=tmax,
                                reduce_lds=rlds)
    print("LDS table bytes", eng.lds_table_bytes, "records pipelines",
          sum(int(d.records) for d in eng._desc_host[:eng.npipes]), flush=True)
    eng.run()
    out, rerun = eng.results()
    assert rerun == []
    for i, (group, plan) in enumerate(pipes):
        ref = _plain(group, plan)
        got = out[i]
        assert torch.equal(got[101].cpu(), ref[101].tensor.cpu()), f"keys of pipeline {i}"
        assert torch.equal(got[302].cpu(), ref[302].tensor.cpu()), f"counts of pipeline {i}"
        np.testing.assert_allclose(got[301].cpu().numpy(), ref[301].tensor.cpu().numpy(),
                                   rtol=1e-12, atol=0, err_msg=f"sums of pipeline {i}")


@pytest.mark.parametrize("fused_min_rows", [1, 50_000])
def test_batched_fused_pipelines_match_plain(device, fused_min_rows):
    """Join pipelines from fused_min_rows probe rows up run as fused groupjoin calls after the
    persistent kernel (1: every pipeline, and no persistent launch); two runs in a row give the
    plain path's results."""
    from tqp_amd import batched, workloads
    probe, build, consts = workloads.config3_sizes(n_pipes=24, lo=1000, hi=300_000)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], device)
             for i in range(len(probe))]
    specs = [batched.match_plan(plan, group) for group, plan in pipes]
    eng = batched.BatchedEngine(specs, tile_rows=2048, fused_min_rows=fused_min_rows)
    print("fused pipelines:", len(eng.fused), "persistent:", eng.npipes, flush=True)
    assert len(eng.fused) == sum(1 for n in probe if n >= fused_min_rows)
    assert eng.npipes == len(probe) - len(eng.fused)
    for rep in range(2):
        eng.run()
        out, rerun = eng.results()
        for i, (group, plan) in enumerate(pipes):
            ref = _plain(group, plan)
            if i in rerun:
                # only a fused call without pairs goes back to the per-operator path
                assert eng.fused and ref[101].tensor.shape[0] == 0, f"pipeline {i} rerun"
                continue
            got = out[i]
            assert torch.equal(got[101].cpu(), ref[101].tensor.cpu()), f"keys of pipeline {i}"
            assert torch.equal(got[302].cpu(), ref[302].tensor.cpu()), f"counts of pipeline {i}"
            np.testing.assert_allclose(got[301].cpu().numpy(), ref[301].tensor.cpu().numpy(),
                                       rtol=1e-12, atol=0, err_msg=f"sums of pipeline {i}")


@pytest.mark.parametrize("fused_min_rows", [None, 10_000])
def test_batched_graph_replay(device, fused_min_rows):
    """The launch sequence captured in a graph replays to the eager result (with fused
    groupjoin calls: their shared presence bytes are cleared inside the graph)."""
    from tqp_amd import batched, workloads
    probe, build, consts = workloads.config3_sizes(n_pipes=8, lo=1000, hi=50_000)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], device)
             for i in range(len(probe))]
    eng = batched.BatchedEngine([batched.match_plan(p, g) for g, p in pipes], tile_rows=2048,
                                fused_min_rows=fused_min_rows)
    print("fused pipelines:", len(eng.fused), flush=True)
    eng.run()
    eager, _ = eng.results()
    eager = {i: {k: v.clone() for k, v in r.items()} for i, r in eager.items()}
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        eng.run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        eng.run()
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        _ = torch.zeros(1000, device=device).sum().item()   # eager work between replays
        got, rerun = eng.results()
        assert rerun == []
        for i in eager:
            assert torch.equal(got[i][101], eager[i][101])
            assert torch.equal(got[i][302], eager[i][302])
            np.testing.assert_allclose(got[i][301].cpu().numpy(), eager[i][301].cpu().numpy(),
                                       rtol=1e-12, atol=0)


def test_batched_rejects_duplicate_build_keys(device):
    from tqp_amd import batched, workloads
    group, plan = workloads.config3_pipeline(0, 20000, 5000, 2556, 0, device)
    group[101].tensor[:10] = 7                     # duplicate selected build keys
    eng = batched.BatchedEngine([batched.match_plan(plan, group)], tile_rows=2048)
    eng.run()
    out, rerun = eng.results()
    assert rerun == [0] and 0 not in out


def test_batched_groupby_pipelines_match_plain(device):
    """Group-by pipelines (filter -> aggregate -> key sort, the config-0 shape) batched together
    with join pipelines give the per-operator results."""
    from tqp_amd import batched, executor, workloads
    base = workloads.to_tensor_group(workloads.config0_tables(n=300_000), device)
    plans = [workloads.config0_plan(ship_day_max=m) for m in (100, 1800, 2556)]
    specs = [batched.match_groupby_plan(p, base) for p in plans]
    assert all(s is not None for s in specs)
    probe, build, consts = workloads.config3_sizes(n_pipes=4, lo=5000, hi=80_000)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], device)
             for i in range(4)]
    specs += [batched.match_plan(p, g) for g, p in pipes]
    eng = batched.BatchedEngine(specs, tile_rows=2048)
    eng.run()
    out, rerun = eng.results()
    assert rerun == []
    for i, plan in enumerate(plans):
        tg = dict(base)
        executor.execute(copy.deepcopy(plan), tg)
        for k in (6, 7, 28):
            assert torch.equal(out[i][k].reshape(tg[k].tensor.shape).cpu(), tg[k].tensor.cpu()), (i, k)
        for k in (22, 23, 24, 25, 26, 27):
            np.testing.assert_allclose(out[i][k].cpu().numpy(), tg[k].tensor.cpu().numpy(),
                                       rtol=1e-12, atol=0, err_msg=f"{i} {k}")


def _host_tables(group):
    """The pipeline's device columns as the oracle's {id: (type, ndarray)} tables."""
    return {cid: (v.tensor_type, v.tensor.cpu().numpy()) for cid, v in group.items()}


def test_config3_full_size_batched_plain_and_oracle(device):
    """The 256 pipelines at their stated sizes (probe rows log-uniform in [10,000, 6,000,000],
    build a quarter of probe, random filter constants): every pipeline's batched result equals
    its per-operator result (keys and counts bit-exact, sums within 1e-12), and eight sampled
    pipelines, the largest among them, equal the oracle's."""
    from oracle import ref_ops as R
    from tqp_amd import batched, workloads
    probe, build, consts = workloads.config3_sizes(n_pipes=256)
    print("config3 full size:", int(probe.sum() + build.sum()), "input rows", flush=True)
    pipes = [workloads.config3_pipeline(i, probe[i], build[i], consts[i][0], consts[i][1], device)
             for i in range(len(probe))]
    specs = [batched.match_plan(plan, group) for group, plan in pipes]
    assert all(s is not None for s in specs)
    eng = batched.BatchedEngine(specs, tile_rows=8192)
    eng.run()
    out, rerun = eng.results()
    assert rerun == []
    for i, (group, plan) in enumerate(pipes):
        ref = _plain(group, plan)
        got = out[i]
        assert torch.equal(got[101].cpu(), ref[101].tensor.cpu()), f"keys of pipeline {i}"
        assert torch.equal(got[302].cpu(), ref[302].tensor.cpu()), f"counts of pipeline {i}"
        np.testing.assert_allclose(got[301].cpu().numpy(), ref[301].tensor.cpu().numpy(),
                                   rtol=1e-12, atol=0, err_msg=f"sums of pipeline {i}")
    print("config3: all 256 batched == per-operator", flush=True)
    g = np.random.default_rng(33)
    sample = sorted({int(np.argmax(probe))} | set(g.choice(len(probe), 7, replace=False).tolist()))
    for i in sample:
        group, plan = pipes[i]
        cpu = R.to_cols(_host_tables(group))
        R.execute(copy.deepcopy(plan), cpu)
        got = out[i]
        print("config3 oracle: pipeline", i, "probe rows", int(probe[i]), flush=True)
        assert torch.equal(got[101].cpu().reshape(cpu[101].tensor.shape), cpu[101].tensor), i
        assert torch.equal(got[302].cpu(), cpu[302].tensor), i
        np.testing.assert_allclose(got[301].cpu().numpy(), cpu[301].tensor.numpy(), rtol=1e-12,
                                   atol=0, err_msg=f"oracle sums of pipeline {i}")
