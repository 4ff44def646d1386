"""GPU tests of the serving loop (mofft_server_* in mofft_hip.h).

Requests of mixed length, batch, precision and direction are admitted
against a device-memory budget and run concurrently on several streams.
The central check: every result is the same alone (a server that runs one
request at a time) and in the mix (many requests admitted while others run,
budget binding) - bit-identical, which is within any tolerance - and each
is within the oracle bar (FP32 1e-6*log2N, FP64 1e-15*log2N relative L2).
Inputs are generated on the device from seeds; the generator's host
restatement (conftest.seeded_device_uniform) regenerates them for the
oracle.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle
from conftest import seeded_device_uniform

pytestmark = pytest.mark.gpu


def tol(n, bits):
    return (1e-6 if bits == 32 else 1e-15) * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def hip_copy():
    import torch
    assert torch.cuda.is_available()
    rt = ctypes.CDLL("libamdhip64.so")
    rt.hipMemcpy.restype = ctypes.c_int
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.restype = ctypes.c_int

    def copy(ptr, n, batch, bits):
        out = np.empty((batch, n), np.complex64 if bits == 32 else np.complex128)
        assert rt.hipDeviceSynchronize() == 0
        assert rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
        return out
    return copy


def mixed_requests(seed, count, kmin, kmax, bmax):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = 1 << int(rng.integers(kmin, kmax + 1))
        b = int(rng.integers(1, bmax + 1))
        bits = 32 if rng.random() < 0.8 else 64
        d = -1 if rng.random() < 0.5 else 1
        reqs.append((n, b, bits, d, int(rng.integers(1, 2**62)), i))
    return reqs


def test_generator_matches_restatement(gpu, hip_copy):
    import torch
    for bits in (32, 64):
        n = 5000
        buf = torch.empty(2 * n, dtype=torch.float32 if bits == 32 else torch.float64, device="cuda")
        gpu.server_generate(buf, bits, n, 12345)
        torch.cuda.synchronize()
        got = buf.cpu().numpy().view(np.complex64 if bits == 32 else np.complex128)
        want = seeded_device_uniform(12345, n, bits)
        assert np.array_equal(got, want)


def test_same_alone_and_in_the_mix(gpu, hip_copy):
    reqs = mixed_requests(77, 40, 8, 17, 6)
    alone = {}
    with gpu.Server(1 << 30, streams=1, max_inflight=1) as srv:
        for (n, b, bits, d, seed, tag) in reqs:
            t = srv.submit(n, b, bits, d, seed=seed, tag=tag)
            (c,) = srv.poll(wait=True)
            assert c.ticket == t and c.status == gpu.SUCCESS
            alone[tag] = hip_copy(c.output, n, b, bits)
            srv.release(t)
        assert srv.stats().peak_inflight == 1
    budget = 24 << 20        # binding: requests wait for each other's memory
    mix = {}
    with gpu.Server(budget, streams=4, max_inflight=16) as srv:
        for (n, b, bits, d, seed, tag) in reqs:
            srv.submit(n, b, bits, d, seed=seed, tag=tag)
        while len(mix) < len(reqs):
            for c in srv.poll(wait=True):
                assert c.status == gpu.SUCCESS
                n, b, bits = reqs[c.tag][:3]
                mix[c.tag] = hip_copy(c.output, n, b, bits)
                srv.release(c.ticket)
        st = srv.stats()
        assert st.completed == len(reqs) and st.pending == 0 and st.inflight == 0
        assert st.peak_bytes_in_use <= budget
        assert st.peak_inflight > 1           # work overlapped on the device
        assert st.bytes_in_use == 0
    for (n, b, bits, d, seed, tag) in reqs:
        assert np.array_equal(alone[tag], mix[tag]), tag
        x = seeded_device_uniform(seed, n * b, bits).reshape(b, n).astype(np.complex128)
        ref = oracle.dft(x, d)
        assert oracle.rel_l2(mix[tag], ref) <= tol(n, bits), (tag, n, b, bits)


def test_budget_and_argument_checks(gpu):
    with gpu.Server(1 << 20) as srv:
        with pytest.raises(gpu.MofftError) as e:
            srv.submit(1 << 16, 4, 64)         # 2 x 4 MiB > 1 MiB budget
        assert e.value.status == gpu.ALLOCATION_FAILURE
        with pytest.raises(gpu.MofftError) as e:
            srv.submit(1 << 10, 65, 32)        # above max_batch
        assert e.value.status == gpu.INVALID_ARGUMENT
        with pytest.raises(gpu.MofftError) as e:
            srv.submit(17, 1, 32)
        assert e.value.status == gpu.UNSUPPORTED_SIZE
        with pytest.raises(gpu.MofftError):
            srv.release(12345)


def test_caller_input_matches_plan(gpu, hip_copy):
    import torch
    n, b = 4096, 3
    x = (np.random.default_rng(4).standard_normal((b, n, 2))).astype(np.float64)
    d_in = torch.from_numpy(x).to("cuda")
    with gpu.Server(64 << 20) as srv:
        t = srv.submit(n, b, 64, gpu.BACKWARD, input=d_in)
        (c,) = srv.poll(wait=True)
        got = hip_copy(c.output, n, b, 64)
        srv.release(t)
    xc = x.view(np.complex128).reshape(b, n)
    y = np.empty_like(xc)
    with gpu.Plan(n, b, 64, gpu.BACKWARD) as p:
        p.execute(xc, y)
    assert np.array_equal(got, y)


def test_poisson_trace(gpu):
    reqs = mixed_requests(5, 300, 8, 20, 16)
    rng = np.random.default_rng(6)
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, len(reqs)))
    with gpu.Server(2 << 30, streams=4, max_inflight=16) as srv:
        cs = srv.run_trace([gpu.Request(n, b, bits, d, None, seed, tag)
                            for (n, b, bits, d, seed, tag) in reqs], arrivals.tolist())
        st = srv.stats()
    assert st.completed == len(reqs) and st.bytes_in_use == 0
    lat = [c.done_seconds - c.submit_seconds for c in cs]
    assert all(c.status == gpu.SUCCESS for c in cs)
    assert all(l > 0 for l in lat)
    assert sorted(c.tag for c in cs) == list(range(len(reqs)))


def test_concurrent_three_pass_requests(gpu, hip_copy):
    """Three-pass schedules need workspace: requests of one plan that run at
    the same time on different streams each get their own from the arena."""
    b = 2
    # the first of these lengths whose tuned FP32 schedule has three passes
    n = next(m for m in (1 << 21, 3 << 20, 1 << 22)
             if gpu.describe(m, b, 32).pass_count == 3)
    reqs = [(n, b, 32, -1 if i % 2 == 0 else 1, 1000 + i, i) for i in range(4)]
    out = {}
    with gpu.Server(1 << 30, streams=4, max_inflight=4) as srv:
        for (nn, bb, bits, d, seed, tag) in reqs:
            srv.submit(nn, bb, bits, d, seed=seed, tag=tag)
        while len(out) < len(reqs):
            for c in srv.poll(wait=True):
                assert c.status == gpu.SUCCESS
                out[c.tag] = hip_copy(c.output, n, b, 32)
                srv.release(c.ticket)
        assert srv.stats().peak_inflight == 4
    for (nn, bb, bits, d, seed, tag) in reqs:
        x = seeded_device_uniform(seed, n * b, 32).reshape(b, n).astype(np.complex128)
        assert oracle.rel_l2(out[tag], oracle.dft(x, d)) <= tol(n, 32), tag


def test_budget_serialises_large_requests(gpu, hip_copy):
    """Requests that do not fit together wait for each other's memory; the
    arena never holds more than the budget."""
    n, b = 1 << 20, 4                  # 64 MiB per FP64 buffer
    with gpu.Server(1 << 30, max_batch=8) as srv:
        srv.submit(n, b, 64, -1, seed=1)
        per = srv.stats().bytes_in_use  # input + output + workspace
        (c,) = srv.poll(wait=True)
        srv.release(c.ticket)
    assert per >= 2 * n * b * 16
    budget = 2 * per + per // 2        # room for two at a time, not three
    reqs = [(n, b, 64, -1, 7 + i, i) for i in range(5)]
    done = {}
    with gpu.Server(budget, streams=4, max_inflight=8, max_batch=8) as srv:
        for (nn, bb, bits, d, seed, tag) in reqs:
            srv.submit(nn, bb, bits, d, seed=seed, tag=tag)
        assert srv.stats().inflight == 2 and srv.stats().pending == 3
        while len(done) < len(reqs):
            for c in srv.poll(wait=True):
                assert c.status == gpu.SUCCESS
                done[c.tag] = hip_copy(c.output, n, b, 64)
                srv.release(c.ticket)
        st = srv.stats()
        assert st.peak_bytes_in_use <= budget and st.peak_inflight == 2
    x = seeded_device_uniform(9, n * b, 64).reshape(b, n)
    assert oracle.rel_l2(done[2], oracle.dft(x, -1)) <= tol(n, 64)


def test_coalesced_requests(gpu, hip_copy):
    """Pending requests of one plan share executions (continuous batching):
    fewer executions than requests, and every result bit-identical to the
    server that runs each request alone, and within the oracle bar."""
    n = 4096
    reqs = [(n, 1 + i % 3, 32, gpu.FORWARD, 500 + i, i) for i in range(24)]
    reqs += [(n, 2, 32, gpu.BACKWARD, 900 + i, 24 + i) for i in range(4)]

    def serve(coalesce):
        out = {}
        with gpu.Server(256 << 20, streams=2, max_inflight=6, max_batch=16,
                        coalesce=coalesce) as srv:
            for (nn, bb, bits, d, seed, tag) in reqs:
                srv.submit(nn, bb, bits, d, seed=seed, tag=tag)
            while len(out) < len(reqs):
                for c in srv.poll(wait=True):
                    assert c.status == gpu.SUCCESS
                    out[c.tag] = hip_copy(c.output, n, reqs[c.tag][1], 32)
                    srv.release(c.ticket)
            st = srv.stats()
            assert st.bytes_in_use == 0 and st.completed == len(reqs)
            assert st.peak_inflight <= 6
        return out, st

    alone, sa = serve(False)
    mixed, sb = serve(True)
    assert sa.executions == len(reqs)
    assert sb.executions < len(reqs) // 2
    for (nn, bb, bits, d, seed, tag) in reqs:
        assert np.array_equal(alone[tag], mixed[tag]), tag
        x = seeded_device_uniform(seed, n * bb, 32).reshape(bb, n).astype(np.complex128)
        assert oracle.rel_l2(mixed[tag], oracle.dft(x, d)) <= tol(n, 32), tag


def test_poll_wait_returns_when_admission_cannot_progress(gpu, hip_copy):
    """A held completion blocks the pending head's memory while nothing is
    in flight: poll(wait=True) returns ALLOCATION_FAILURE instead of waiting
    forever; after the release the head is admitted and completes."""
    n, b = 1 << 16, 4
    with gpu.Server(1 << 30, max_batch=4) as srv:
        srv.submit(n, b, 64, -1, seed=1)
        per = srv.stats().bytes_in_use
        (c,) = srv.poll(wait=True)
        srv.release(c.ticket)
    with gpu.Server(per + per // 2, max_batch=4) as srv:
        ta = srv.submit(n, b, 64, -1, seed=1, tag=1)
        (ca,) = srv.poll(wait=True)
        assert ca.ticket == ta and ca.status == gpu.SUCCESS
        tb = srv.submit(n, b, 64, 1, seed=2, tag=2)   # needs A's memory
        with pytest.raises(gpu.MofftError) as e:
            srv.poll(wait=True)
        assert e.value.status == gpu.ALLOCATION_FAILURE
        srv.release(ta)
        (cb,) = srv.poll(wait=True)
        assert cb.ticket == tb and cb.status == gpu.SUCCESS
        got = hip_copy(cb.output, n, b, 64)
        srv.release(tb)
    x = seeded_device_uniform(2, n * b, 64).reshape(b, n)
    assert oracle.rel_l2(got, oracle.dft(x, 1)) <= tol(n, 64)


def test_run_trace_fails_a_head_blocked_by_held_memory(gpu):
    """run_trace with completions still held from before the call: a head
    that cannot fit fails with ALLOCATION_FAILURE instead of stalling the
    replay; the others complete."""
    n, b = 1 << 16, 4
    with gpu.Server(1 << 30, max_batch=4) as srv:
        srv.submit(n, b, 64, -1, seed=1)
        per = srv.stats().bytes_in_use
        (c,) = srv.poll(wait=True)
        srv.release(c.ticket)
    with gpu.Server(per + per // 2, max_batch=4) as srv:
        srv.submit(n, b, 64, -1, seed=1)
        (held,) = srv.poll(wait=True)
        reqs = [gpu.Request(n, b, 64, -1, None, 5, 0), gpu.Request(256, 1, 32, -1, None, 6, 1)]
        cs = srv.run_trace(reqs, [0.0, 0.0])
        assert cs[0].status == gpu.ALLOCATION_FAILURE
        assert cs[1].status == gpu.SUCCESS


def _copy_transform(ptr, n, t, bits):
    """Transform t (n points) of a device result, to the host."""
    rt = ctypes.CDLL("libamdhip64.so")
    rt.hipMemcpy.restype = ctypes.c_int
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(n, np.complex64 if bits == 32 else np.complex128)
    assert rt.hipMemcpy(out.ctypes.data, ptr + t * out.nbytes, out.nbytes, 2) == 0
    return out


def c5_requests(seed, count):
    """Requests from the C5 serving distribution: N = 2^U{8..24}, batch
    U{1..64}, 80 % FP32 / 20 % FP64, half forward, seeded device inputs."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = 1 << int(rng.integers(8, 25))
        b = int(rng.integers(1, 65))
        bits = 32 if rng.random() < 0.8 else 64
        d = -1 if rng.random() < 0.5 else 1
        reqs.append((n, b, bits, d, int(rng.integers(1, 2**62)), i))
    return reqs


def test_c5_distribution_full_size(gpu, hip_copy):
    """The C5 serving configuration at its stated sizes: 96 requests drawn
    from its distribution plus its largest request, FP64 N = 2^24 x 64 (16
    GiB in, 16 GiB out), submitted together under the 64 GB admission budget
    so that admission binds (the large request waits for memory and holds
    others back).  Every request completes successfully; for each length
    N >= 2^21 one sampled transform of one request is checked against the
    oracle at the precision bar; a subset of the shorter requests is served
    again alone (one request at a time) and must be bit-identical to its
    result in the mix."""
    reqs = c5_requests(2000, 96)
    big = (1 << 24, 64, 64, -1, 424242, len(reqs))
    reqs.insert(10, big)
    reqs = [(n, b, bits, d, seed, i) for i, (n, b, bits, d, seed, _) in enumerate(reqs)]
    budget = 64 << 30
    subset = [r for r in reqs if r[0] <= 1 << 16][:12]
    sub_tags = {r[5] for r in subset}
    checks = {}
    for r in reqs:
        if r[0] >= 1 << 21 and r[0] not in checks:
            checks[r[0]] = r[5]
    assert len(checks) == 4 and subset
    mix, sampled = {}, {}
    done = 0
    with gpu.Server(budget, streams=4, max_inflight=16) as srv:
        for (n, b, bits, d, seed, tag) in reqs:
            srv.submit(n, b, bits, d, seed=seed, tag=tag)
        st = srv.stats()
        assert st.pending > 0          # admission binds: requests wait for memory
        while done < len(reqs):
            for c in srv.poll(wait=True):
                assert c.status == gpu.SUCCESS, c.tag
                n, b, bits = reqs[c.tag][:3]
                if c.tag in sub_tags:
                    mix[c.tag] = hip_copy(c.output, n, b, bits)
                if c.tag in checks.values():
                    t = int(np.random.default_rng(c.tag).integers(0, b))
                    sampled[c.tag] = (t, _copy_transform(c.output, n, t, bits))
                srv.release(c.ticket)
                done += 1
        st = srv.stats()
        assert st.completed == len(reqs) and st.bytes_in_use == 0
        assert st.peak_bytes_in_use <= budget
        assert st.peak_bytes_in_use >= 2 * (1 << 24) * 64 * 16   # the big one ran
    for n, tag in sorted(checks.items()):
        _, b, bits, d, seed, _ = reqs[tag]
        t, y = sampled[tag]
        x = seeded_device_uniform(seed, n, bits, start=t * n).astype(np.complex128)
        err = oracle.rel_l2(y, oracle.fft_dd(x[None, :], d)[0])
        print(f"C5 check N=2^{n.bit_length() - 1} batch {b} FP{bits} transform {t}: rel-L2 {err:.2e}")
        assert err <= tol(n, bits), (n, bits, err)
    alone = {}
    with gpu.Server(1 << 30, streams=1, max_inflight=1) as srv:
        for (n, b, bits, d, seed, tag) in subset:
            t = srv.submit(n, b, bits, d, seed=seed, tag=tag)
            (c,) = srv.poll(wait=True)
            assert c.ticket == t and c.status == gpu.SUCCESS
            alone[tag] = hip_copy(c.output, n, b, bits)
            srv.release(t)
    for tag in sub_tags:
        assert np.array_equal(alone[tag], mix[tag]), tag


def test_coalescing_flag_rejects_unknown_bits(gpu):
    import ctypes as C
    o = gpu.ServerOptions()
    gpu.lib().mofft_server_options_init(C.byref(o))
    o.flags = 4
    h = C.c_void_p()
    assert gpu.lib().mofft_server_create(C.byref(h), C.byref(o)) == gpu.INVALID_ARGUMENT
