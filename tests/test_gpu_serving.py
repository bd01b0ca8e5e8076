"""Continuous-batching server: results equal the standalone solve bit for bit,
alone and in a mix; requests overlap on the device; budget admission."""
import ctypes
import threading

import numpy as np
import pytest

import xrtk
from xrtk import workloads as W

pytestmark = pytest.mark.gpu


def grid(scr):
    return xrtk.WaveParams(scr.center, scr.dx, scr.dz, scr.nx, scr.nz)


@pytest.fixture(scope="module")
def sets(ctx):
    specs = W.batch_instances(n=8, seed=31, max_rays=120000, max_side=160)
    rsets = [xrtk.RaySet(ctx, W.instance_rays(s)) for s in specs]
    yield specs, rsets
    for r in rsets:
        r.release()


def jobs_of(specs, rsets, energies=None):
    out = []
    for i, (s, r) in enumerate(zip(specs, rsets)):
        e = s["energy"] if energies is None else energies[i]
        out.append(xrtk.Job(r, e, grid(W.instance_screen(s))))
    return out


def test_server_results_equal_standalone(ctx, sets):
    """With a quantum long enough for the configuration's own ray chunks, the
    server's plan is the batched solve's: the same bits."""
    specs, rsets = sets
    jobs = jobs_of(specs, rsets)
    alone = [xrtk.solve_jobs(ctx, [j])[0] for j in jobs]
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, max_active=64, quantum_ms=100.0)
    try:
        tickets = [srv.submit(j) for j in jobs]
        for t, a in zip(tickets, alone):
            got, ts = srv.wait(t, timeout_ms=60000)
            assert np.array_equal(got, a)
            assert ts[0] <= ts[1] <= ts[2] <= ts[3]
        st = srv.stats()
        assert st["completed"] == len(jobs) and st["max_concurrent"] > 1
    finally:
        srv.close()


@pytest.mark.parametrize("persistent", [0, 1])
def test_server_launch_shapes(ctx, sets, persistent):
    """The server on a given launch shape (P = 16, one wave per group;
    one group per work item, or persistent groups on the queue) equals the
    batched solve on the same shape bit for bit, with every request's items
    run (more items per launch than resident groups)."""
    specs, rsets = sets
    big_spec = {"n_rays": 1 << 20, "side": 512, "energy": 10000.0, "distance": 1000.0,
                "seed": 78, "half": 0.2 * 0.1 * (1000.0 + W.SRC_DIST) / W.SRC_DIST}
    big_rs = xrtk.RaySet(ctx, W.instance_rays(big_spec))
    try:
        # the 1 M x 512^2 request has 32768 work items (128 chunks x 256 pixel
        # blocks): a 100 ms quantum holds ~2e4 of them, against 8192 groups
        jobs = jobs_of(specs, rsets) + [
            xrtk.Job(big_rs, big_spec["energy"], grid(W.instance_screen(big_spec)))]
        shape = dict(pixels_per_lane=16, waves_per_group=1, persistent=persistent)
        alone = [xrtk.solve_jobs(ctx, [j], **shape)[0] for j in jobs]
        srv = xrtk.Server(ctx, hbm_budget_bytes=16 << 30, quantum_ms=100.0, **shape)
        try:
            got = [srv.wait(t, timeout_ms=60000)[0] for t in [srv.submit(j) for j in jobs]]
            st = srv.stats()
        finally:
            srv.close()
        for k, (g, a) in enumerate(zip(got, alone)):
            assert np.array_equal(g, a), k
        assert st["completed"] == len(jobs)
    finally:
        big_rs.release()


def test_short_quantum_chunks_match_standalone(ctx, sets):
    """A short quantum cuts the requests' ray chunks (one work item stays
    within half a quantum): the FP32 partial sums change, the result stays
    within FP32 rounding of the batched solve."""
    specs, rsets = sets
    jobs = jobs_of(specs, rsets)
    alone = [xrtk.solve_jobs(ctx, [j])[0] for j in jobs]
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, max_active=64, quantum_ms=0.5)
    try:
        got = [srv.wait(t, timeout_ms=60000)[0] for t in [srv.submit(j) for j in jobs]]
    finally:
        srv.close()
    for g, a in zip(got, alone):
        assert W.rel_l2(g, a) < 2e-6


@pytest.mark.parametrize("policy", ["shortest", "arrival"])
def test_alone_and_in_mix_identical(ctx, sets, policy):
    """The same request solved alone and while others arrive and run."""
    specs, rsets = sets
    jobs = jobs_of(specs, rsets)
    target = jobs[2]
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, max_active=64, quantum_ms=1.0,
                      policy=policy)
    try:
        alone, _ = srv.wait(srv.submit(target), timeout_ms=60000)
        tickets = []
        lock = threading.Lock()

        def feeder(js):
            for j in js:
                t = srv.submit(j)
                with lock:
                    tickets.append(t)

        th = [threading.Thread(target=feeder, args=(jobs[k::2],)) for k in range(2)]
        for t in th:
            t.start()
        mixed_t = srv.submit(target)
        for t in th:
            t.join()
        mixed, _ = srv.wait(mixed_t, timeout_ms=60000)
        for t in tickets:
            srv.wait(t, timeout_ms=60000)
        assert np.array_equal(alone, mixed)
        assert srv.stats()["max_concurrent"] > 1
    finally:
        srv.close()


def test_energy_scan_reuses_resident_rays(ctx, sets):
    specs, rsets = sets
    s, r = specs[0], rsets[0]
    energies = [8000.0, 9000.0, 12000.0]
    # a quantum long enough for the configuration's own ray chunks: the
    # batched solve's plan, so the same bits
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, quantum_ms=100.0)
    try:
        ts = [srv.submit(xrtk.Job(r, e, grid(W.instance_screen(s)))) for e in energies]
        outs = [srv.wait(t, timeout_ms=60000)[0] for t in ts]
    finally:
        srv.close()
    for e, o in zip(energies, outs):
        ref = xrtk.solve_jobs(ctx, [xrtk.Job(r, e, grid(W.instance_screen(s)))])[0]
        assert np.array_equal(o, ref)


def test_budget_rejects_and_queues(ctx, sets):
    specs, rsets = sets
    jobs = jobs_of(specs, rsets)
    srv = xrtk.Server(ctx, hbm_budget_bytes=1 << 20)  # 1 MiB: nothing fits
    try:
        with pytest.raises(xrtk.XrtkError) as ei:
            srv.submit(jobs[0])
        assert ei.value.code == xrtk._lib.ENOMEM
        assert srv.stats()["rejected"] == 1
    finally:
        srv.close()
    # a budget that holds one request at a time still completes them all
    big = max(jobs, key=lambda j: j.rays.n * j.screen.nx * j.screen.nz)
    srv = xrtk.Server(ctx, hbm_budget_bytes=3 << 30, max_active=1)
    try:
        ts = [srv.submit(j) for j in jobs]
        for t in ts:
            srv.wait(t, timeout_ms=120000)
        st = srv.stats()
        assert st["completed"] == len(jobs) and st["max_concurrent"] == 1
    finally:
        srv.close()
    assert big is not None


def test_release_of_busy_rayset_refused(ctx, sets):
    specs, rsets = sets
    rays = W.instance_rays(specs[1])
    rs = xrtk.RaySet(ctx, rays)
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30)
    try:
        t = srv.submit(xrtk.Job(rs, specs[1]["energy"], grid(W.instance_screen(specs[1]))))
        with pytest.raises(xrtk.XrtkError) as ei:
            rs.release()
        assert ei.value.code in (xrtk._lib.EAGAIN,)
        srv.wait(t, timeout_ms=60000)
    finally:
        srv.close()
    rs.release()


def test_shortest_first_lets_small_requests_pass(ctx, sets):
    """Shortest-remaining-first: small requests that arrive while a large
    one runs finish before it; every result still matches the standalone
    solve (to FP32 rounding: the short quantum cuts the ray chunks)."""
    specs, rsets = sets
    big_spec = {"n_rays": 1 << 20, "side": 512, "energy": 10000.0, "distance": 1000.0,
                "seed": 77, "half": 0.2 * 0.1 * (1000.0 + W.SRC_DIST) / W.SRC_DIST}
    big_rs = xrtk.RaySet(ctx, W.instance_rays(big_spec))
    try:
        big = xrtk.Job(big_rs, big_spec["energy"], grid(W.instance_screen(big_spec)))
        small = sorted(jobs_of(specs, rsets), key=lambda j: j.rays.n * j.screen.nx * j.screen.nz)[:4]
        ref = [xrtk.solve_jobs(ctx, [j])[0] for j in small]
        srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, quantum_ms=2.0, policy="shortest")
        try:
            tb = srv.submit(big)
            ts = [srv.submit(j) for j in small]
            outs = [srv.wait(t, timeout_ms=60000) for t in ts]
            _, tbig = srv.wait(tb, timeout_ms=60000)
        finally:
            srv.close()
        for (got, t), r in zip(outs, ref):
            assert W.rel_l2(got, r) < 2e-6
            assert t[3] < tbig[3]
    finally:
        big_rs.release()


def test_invalid_policy_rejected(ctx):
    cfg = xrtk._lib.XrtkServerCfg(1 << 30, 4, 2.0, 0, 7)
    h = ctypes.c_void_p()
    rc = xrtk._lib.lib.xrtk_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h))
    assert rc == xrtk._lib.EINVAL and not h.value


@pytest.mark.parametrize("policy", ["shortest", "arrival"])
def test_admission_follows_policy(ctx, sets, policy):
    """One active request at a time: while the largest runs, a medium and
    then a small request queue.  Shortest-remaining-first admits the small
    one first, arrival order the medium one; both equal the standalone
    solve."""
    specs, rsets = sets
    jobs = sorted(jobs_of(specs, rsets), key=lambda j: j.rays.n * j.screen.nx * j.screen.nz)
    small, medium, big = jobs[0], jobs[len(jobs) // 2], jobs[-1]
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, max_active=1, quantum_ms=100.0,
                      policy=policy)
    try:
        tb = srv.submit(big)
        tm = srv.submit(medium)
        tsm = srv.submit(small)
        outs = {k: srv.wait(t, timeout_ms=60000) for k, t in
                (("big", tb), ("medium", tm), ("small", tsm))}
        assert srv.stats()["max_concurrent"] == 1
    finally:
        srv.close()
    for k, j in (("big", big), ("medium", medium), ("small", small)):
        assert np.array_equal(outs[k][0], xrtk.solve_jobs(ctx, [j])[0]), k
    t_small, t_medium = outs["small"][1], outs["medium"][1]
    if policy == "shortest":
        assert t_small[1] < t_medium[1] and t_small[3] < t_medium[3]
    else:
        assert t_medium[1] < t_small[1] and t_medium[3] < t_small[3]


def test_non_square_screens_equal_batched(ctx, sets):
    """Requests on non-square and one-pixel-wide screens (edge tiles padded
    along one or both axes, the tile shape chosen per screen) equal the
    batched solve bit for bit, and the batched solve agrees with the FP64
    kernel on the same jobs."""
    specs, rsets = sets
    s, r = specs[0], rsets[0]
    screens = [xrtk.WaveParams((0.001, s["distance"], -0.002), 0.02, 0.015, nx, nz)
               for nx, nz in ((300, 70), (33, 100), (1, 257), (257, 1))]
    jobs = [xrtk.Job(r, s["energy"], w) for w in screens]
    batched = xrtk.solve_jobs(ctx, jobs)
    f64 = xrtk.solve_jobs(ctx, jobs, mode="f64")
    srv = xrtk.Server(ctx, hbm_budget_bytes=8 << 30, quantum_ms=100.0)
    try:
        got = [srv.wait(t, timeout_ms=60000)[0] for t in [srv.submit(j) for j in jobs]]
    finally:
        srv.close()
    for k, (g, b, d) in enumerate(zip(got, batched, f64)):
        assert np.array_equal(g, b), k
        # (the FP64 kernel's own phase rounding reaches ~5e-5 at 20 m)
        assert W.rel_l2(b, d) < 1e-4, k


def test_poisson_energy_scan_stream(ctx, sets):
    """A served stream in the form of the energy-scan workload, shortened:
    Poisson arrivals (1000 req/s for 0.3 s, seeded: dense enough that these
    small requests share launches) of energy scans (1-50 energies, 5 eV
    apart) over the resident ray sets, admitted under a budget and batched
    continuously.  Every result equals the batched solve
    of the same (ray set, energy, screen) to FP32 rounding (the 10 ms
    quantum cuts the ray chunks), requests overlap on the device, and the
    timestamps are ordered submit <= admit <= first launch <= done."""
    import time
    specs, rsets = sets
    rng = np.random.default_rng(17)
    reqs = []
    while len(reqs) < 300:
        i = int(rng.integers(len(specs)))
        m = int(rng.integers(1, 51))
        reqs += [(i, specs[i]["energy"] + 5.0 * (k - m // 2)) for k in range(m)]
    arrivals = np.cumsum(rng.exponential(1.0 / 1000.0, len(reqs)))
    n = int(np.searchsorted(arrivals, 0.3))
    reqs, arrivals = reqs[:n], arrivals[:n]
    screens = [grid(W.instance_screen(s)) for s in specs]
    srv = xrtk.Server(ctx, hbm_budget_bytes=4 << 30, max_active=64, quantum_ms=10.0)
    try:
        t0 = time.perf_counter()
        tickets = []
        for (i, e), ta in zip(reqs, arrivals):
            d = t0 + ta - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append(srv.submit(xrtk.Job(rsets[i], e, screens[i])))
        got = [srv.wait(t, timeout_ms=120000) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    lat = np.array([ts[3] - ts[0] for _, ts in got]) * 1e3
    print(f"{n} requests in {arrivals[-1]:.2f} s: p50 {np.percentile(lat, 50):.1f} ms, "
          f"p99 {np.percentile(lat, 99):.1f} ms, max concurrent {st['max_concurrent']}")
    assert st["completed"] == n and st["max_concurrent"] > 1
    for (out, ts), (i, e) in zip(got, reqs):
        assert ts[0] <= ts[1] <= ts[2] <= ts[3]
    # every distinct (set, energy) against the batched solve
    seen = {}
    for (out, _), (i, e) in zip(got, reqs):
        if (i, e) not in seen:
            (seen[(i, e)],) = xrtk.solve_jobs(ctx, [xrtk.Job(rsets[i], e, screens[i])])
        assert W.rel_l2(out, seen[(i, e)]) < 2e-6
