"""GPU tests of the persistent uneven batch (mofft_batch_* in mofft_hip.h).

Many transforms of different power-of-two lengths, directions and
placements run in one persistent launch.  Every numeric check is against
the FP64 oracle with the path's bars (FP32 1e-6*log2N, FP64 1e-15*log2N
relative L2, per transform); identity checks compare the batch with the
single-transform plan of the same chain and pass grouping (same arithmetic,
other tile shapes) and graph replays with direct launches.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import seeded_complex

pytestmark = pytest.mark.gpu


def tol(n, bits):
    return (1e-6 if bits == 32 else 1e-15) * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available()
    return torch


def cdtype(bits):
    return np.complex64 if bits == 32 else np.complex128


def make_jobs(torch, seed, count, kmin, kmax, bits, dist="normal"):
    """Seeded lengths 2^U{kmin..kmax}, directions, placements and inputs."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(kmin, kmax + 1, count)
    lengths = [1 << int(k) for k in ks]
    dirs = [-1 if d else 1 for d in rng.integers(0, 2, count)]
    inplace = [bool(p) for p in rng.integers(0, 2, count)]
    xs, ins, outs = [], [], []
    for i, n in enumerate(lengths):
        x = seeded_complex((n,), seed=seed * 100003 + i, dist=dist, dtype=cdtype(bits))
        d_in = torch.from_numpy(x).to("cuda")
        d_out = d_in if inplace[i] else torch.empty_like(d_in)
        xs.append(x)
        ins.append(d_in)
        outs.append(d_out)
    return lengths, dirs, inplace, xs, ins, outs


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("graph,share_all", [(True, False), (False, False), (True, True)])
def test_uneven_batch_vs_oracle(gpu, torch_dev, bits, graph, share_all):
    """Owned transforms (one workgroup each) and shared ones (pass tiles
    over many workgroups, counter-synchronised), graph replay and direct
    launch."""
    torch = torch_dev
    lengths, dirs, inplace, xs, ins, outs = make_jobs(torch, 11 + bits, 48, 1, 18, bits)
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits, graph=graph, share_all=share_all) as bp:
        d = bp.description()
        assert d.job_count == 48 and d.graph == int(graph)
        assert d.multi_pass_jobs == sum(n > 4096 for n in lengths)
        if share_all:
            assert d.shared_jobs == d.multi_pass_jobs
        bp.execute()
        torch.cuda.synchronize()
    for i, n in enumerate(lengths):
        y = outs[i].cpu().numpy()
        ref = oracle.dft(xs[i].astype(np.complex128), dirs[i])
        err = oracle.rel_l2(y, ref)
        assert err <= tol(n, bits), (i, n, dirs[i], inplace[i], err)


@pytest.mark.parametrize("bits", [32, 64])
def test_batch_equals_single_plans(gpu, torch_dev, bits):
    """Same result alone (one single-transform plan of the same chain and
    pass grouping) as in the batch: bit-identical."""
    torch = torch_dev
    lengths, dirs, inplace, xs, ins, outs = make_jobs(torch, 5 + bits, 24, 4, 18, bits)
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits) as bp:
        bp.execute()
    torch.cuda.synchronize()
    for i, n in enumerate(lengths):
        desc = gpu.batch_describe_length(n, bits)
        x = xs[i]
        y = np.empty_like(x)
        with gpu.Plan(n, 1, bits, dirs[i], radices=desc.chain, pass_mask=desc.pass_mask) as p:
            assert p.description().passes == desc.passes
            p.execute(x, y)
        assert np.array_equal(outs[i].cpu().numpy(), y), (i, n)


def test_full_uneven_batch_round_trip(gpu, torch_dev):
    """4096 FP32 transforms, N = 2^U{6..18}, 50/50 directions, mixed
    placement: forward then backward batches return N*x for every
    transform; a seeded sample is checked against the oracle."""
    torch = torch_dev
    bits = 32
    lengths, dirs, inplace, xs, ins, outs = make_jobs(torch, 2026, 4096, 6, 18, bits)
    keep = [t.clone() for t in ins]
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits) as bp:
        d = bp.description()
        assert d.job_count == 4096 and d.graph == 1
        bp.execute()
    torch.cuda.synchronize()
    rng = np.random.default_rng(7)
    for i in rng.choice(4096, 24, replace=False):
        n = lengths[i]
        ref = oracle.dft(xs[i].astype(np.complex128), dirs[i])
        assert oracle.rel_l2(outs[i].cpu().numpy(), ref) <= tol(n, bits), (i, n)
    # inverse batch on the outputs (opposite directions, same placements)
    back = [o if ip else torch.empty_like(o) for o, ip in zip(outs, inplace)]
    with gpu.BatchPlan(lengths, [-d for d in dirs], outs, back, bits) as inv:
        inv.execute()
    torch.cuda.synchronize()
    worst = 0.0
    for i, n in enumerate(lengths):
        r = back[i] / n
        e = (torch.linalg.vector_norm(r - keep[i]) / torch.linalg.vector_norm(keep[i])).item()
        worst = max(worst, e / tol(n, bits))
    assert worst <= 1.0, worst


def test_graph_replays_match_eager(gpu, torch_dev):
    """Replays of the captured graph, with eager work and synchronisation
    between them, equal a direct launch every time."""
    torch = torch_dev
    bits = 32
    lengths, dirs, inplace, xs, ins, outs = make_jobs(torch, 99, 64, 2, 17, bits)
    keep = [t.clone() for t in ins]
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits, graph=False) as direct:
        direct.execute()
        want = [o.clone() for o in outs]
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits, graph=True) as bp:
        stream = torch.cuda.Stream()
        for rep in range(4):
            for a, b in zip(ins, keep):
                a.copy_(b)
            junk = torch.randn(1 << 20, device="cuda").sum()   # eager work
            torch.cuda.synchronize()
            bp.execute_async(stream.cuda_stream)
            stream.synchronize()
            bp.check()
            for i in range(len(lengths)):
                assert torch.equal(outs[i], want[i]), (rep, i)
            del junk


def test_rebind_buffers(gpu, torch_dev):
    torch = torch_dev
    bits = 64
    lengths, dirs, inplace, xs, ins, outs = make_jobs(torch, 3, 16, 3, 16, bits)
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits) as bp:
        ins2 = [t.flip(0).contiguous() for t in ins]
        outs2 = [a if ip else torch.empty_like(a) for a, ip in zip(ins2, inplace)]
        bp.set_buffers(ins2, outs2)
        bp.execute()
        # placement must not change
        swapped = [a if not ip else torch.empty_like(a) for a, ip in zip(ins2, inplace)]
        with pytest.raises(gpu.MofftError) as e:
            bp.set_buffers(ins2, swapped)
        assert e.value.status == gpu.INVALID_ARGUMENT
    torch.cuda.synchronize()
    for i, n in enumerate(lengths):
        ref = oracle.dft(xs[i][::-1].astype(np.complex128), dirs[i])
        assert oracle.rel_l2(outs2[i].cpu().numpy(), ref) <= tol(n, bits), i


def test_dependency_wait_limit_reports_error(gpu, torch_dev):
    """Force the bounded dependency wait to give up (a spin limit of one
    poll; half of the 128 workgroups start on second-pass tiles while the
    first pass is still running): the launch
    ends normally and the execution reports MOFFT_INTERNAL_ERROR.  A plan
    with the default limit on the same buffers then succeeds."""
    torch = torch_dev
    n = 1 << 18
    x = seeded_complex((n,), seed=1, dtype=np.complex64)
    d_in = torch.from_numpy(x).to("cuda")
    d_out = torch.empty_like(d_in)
    with gpu.BatchPlan([n], [-1], [d_in], [d_out], 32, graph=False, grid=128,
                       spin_limit=1) as bp:
        assert bp.description().shared_jobs == 1   # a lone transform is shared
        with pytest.raises(gpu.MofftError) as e:
            bp.execute()
        assert e.value.status == gpu.INTERNAL_ERROR
    with gpu.BatchPlan([n], [-1], [d_in], [d_out], 32) as bp:
        bp.execute()
    torch.cuda.synchronize()
    assert oracle.rel_l2(d_out.cpu().numpy(), oracle.dft(x.astype(np.complex128), -1)) <= tol(n, 32)


def test_batch_rejects_host_pointers(gpu):
    x = np.zeros(64, np.complex64)
    y = np.zeros(64, np.complex64)
    with pytest.raises(gpu.MofftError) as e:
        gpu.BatchPlan([64], [-1], [x], [y], 32)
    assert e.value.status == gpu.INVALID_ARGUMENT


@pytest.mark.parametrize("bits", [32, 64])
def test_three_pass_and_edge_shapes(gpu, torch_dev, bits):
    """Transforms of three strided passes (2^19..2^21), in and out of place,
    next to the smallest lengths (2, 4), one of each, and a lone transform."""
    torch = torch_dev
    lengths = [1 << 19, 1 << 20, 1 << 21, 2, 4, 1 << 21, 8]
    dirs = [-1, 1, -1, 1, -1, 1, -1]
    inplace = [True, False, True, False, True, False, False]
    rng = np.random.default_rng(bits)
    xs, ins, outs = [], [], []
    for n, ip in zip(lengths, inplace):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(cdtype(bits))
        d = torch.from_numpy(x).to("cuda")
        xs.append(x)
        ins.append(d)
        outs.append(d if ip else torch.empty_like(d))
    for n in lengths[:3]:
        assert gpu.batch_describe_length(n, bits).pass_count == 3
    with gpu.BatchPlan(lengths, dirs, ins, outs, bits) as bp:
        assert bp.description().multi_pass_jobs == 4
        bp.execute()
    torch.cuda.synchronize()
    for i, n in enumerate(lengths):
        ref = oracle.dft(xs[i].astype(np.complex128), dirs[i])
        assert oracle.rel_l2(outs[i].cpu().numpy(), ref) <= tol(n, bits), (i, n)
    # a lone two-element transform
    x = np.array([1 + 2j, 3 - 1j], dtype=cdtype(bits))
    d = torch.from_numpy(x).to("cuda")
    with gpu.BatchPlan([2], [-1], [d], [d], bits) as bp:
        bp.execute()
    torch.cuda.synchronize()
    assert np.allclose(d.cpu().numpy(), [4 + 1j, -2 + 3j])


def test_many_equal_small_transforms(gpu, torch_dev):
    """3000 transforms of one length: a single gathered segment."""
    torch = torch_dev
    n, count = 64, 3000
    x = seeded_complex((count, n), seed=64, dtype=np.complex64)
    d_in = torch.from_numpy(x).to("cuda")
    d_out = torch.empty_like(d_in)
    ins = [d_in[i] for i in range(count)]
    outs = [d_out[i] for i in range(count)]
    with gpu.BatchPlan([n] * count, [-1] * count, ins, outs, 32) as bp:
        d = bp.description()
        assert d.distinct_lengths == 1 and d.segment_count == 1 and d.multi_pass_jobs == 0
        bp.execute()
    torch.cuda.synchronize()
    ref = oracle.dft(x.astype(np.complex128), -1)
    assert oracle.rel_l2(d_out.cpu().numpy(), ref) <= tol(n, 32)
