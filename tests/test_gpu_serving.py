"""Served mode (continuous batching): requests submitted while others are on
the GPU are admitted at the next SCF-iteration boundary, retire at their own
convergence, and give the same result as when solved alone (total energy
within 1e-9 eV, same iteration count, density within 1e-10)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.0 / 0.03674903


def _job(kind, size, seed, precision="fp64", rms=1e-8, damping=0.7):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    Z, x = W.alkane(size, seed=seed) if kind == "alkane" else W.water_cluster(size, seed=seed, jitter=0.02)
    return Job("pm3", Z, W.to_bohr(x), precision=precision, rms_density=rms, diis_num_error_vect=5,
               diis_start_error=0.1, diis_end_error=0.0, damping_weight=damping, damping_thresh=0.02, max_iter=200)


def test_served_matches_alone_and_overlaps(engine):
    print("served overlap", flush=True)
    # the big request needs far more SCF iterations than the small later ones
    big = _job("water", 64, 31, rms=1e-10)
    first = [big, _job("alkane", 9, 32), _job("water", 12, 33, "mixed")]
    second = [_job("alkane", 1, 34, damping=0.0), _job("alkane", 2, 35, "mixed", damping=0.0),
              _job("water", 2, 36, damping=0.0)]
    t_sub, ids = {}, {}
    for j in first:
        ids[id(j)] = engine.submit(j)
        t_sub[id(j)] = time.perf_counter()
    time.sleep(0.005)
    for j in second:
        ids[id(j)] = engine.submit(j)
        t_sub[id(j)] = time.perf_counter()
    res = {}
    for j in first + second:
        st = engine.wait(ids[id(j)], 60.0)
        assert st == 0, st
        res[id(j)] = engine.get_result(ids[id(j)], j, want_P=True)
        assert engine.release(ids[id(j)]) == 0
    done = {k: t_sub[k] + r["latency_s"] for k, r in res.items()}
    # the small requests submitted after the big one joined it on the GPU and
    # retired at their own convergence, before it
    assert max(r["iterations"] for r in (res[id(j)] for j in second)) + 3 < res[id(big)]["iterations"]
    assert max(done[id(j)] for j in second) < done[id(big)]
    for j in first + second:
        r = res[id(j)]
        alone = engine.scf_batch([j], want_P=True)[0]
        assert r["status"] == 0 and alone["status"] == 0
        assert abs(r["energy"] - alone["energy"]) * EV <= 1e-9
        assert r["iterations"] == alone["iterations"]
        assert np.max(np.abs(r["P"] - alone["P"])) <= 1e-10
        assert r["latency_s"] > 0.0


def test_served_graph_replay_after_smaller_launches(engine):
    """A captured SCF iteration keeps replaying correctly after another
    cohort launched the same eigensolver kernels with less dynamic LDS: the
    kernels' LDS limit only ever grows, so the recorded launch stays valid."""
    print("served replay after smaller launches", flush=True)
    big = _job("water", 64, 41, rms=1e-10)          # 384 AOs, device-memory eigensolver class
    small = [_job("water", 25, 42), _job("alkane", 24, 43)]   # 150 and 146 AOs
    g0 = engine.graph_stats()
    jb = engine.submit(big)
    t0 = time.perf_counter()
    while engine.graph_stats()["captured"] == g0["captured"] and time.perf_counter() - t0 < 20.0:
        time.sleep(0.001)   # until the big cohort has captured its iteration
    ids = [engine.submit(j) for j in small]
    res = {}
    for k, j in [(jb, big)] + list(zip(ids, small)):
        st = engine.wait(k, 60.0)
        assert st == 0, st
        res[k] = engine.get_result(k, j, want_P=True)
        assert engine.release(k) == 0
    g1 = engine.graph_stats()
    assert g1["captured"] > g0["captured"] and g1["replays"] > g0["replays"]
    for k, j in [(jb, big)] + list(zip(ids, small)):
        alone = engine.scf_batch([j], want_P=True)[0]
        assert res[k]["status"] == 0 and alone["status"] == 0
        assert abs(res[k]["energy"] - alone["energy"]) * EV <= 1e-9
        assert res[k]["iterations"] == alone["iterations"]
        assert np.max(np.abs(res[k]["P"] - alone["P"])) <= 1e-10


def test_served_results_taken_while_cohorts_capture(engine):
    """Results fetched and released by the caller while later cohorts start
    and capture their iteration graphs: every request succeeds, and the
    results equal the alone solves (the caller's copies and frees never touch
    a cohort stream)."""
    print("served results under capture", flush=True)
    waves = [[_job("water", 3 + (7 * w + k) % 20, 60 + 10 * w + k) if k % 2 else
              _job("alkane", 2 + (5 * w + k) % 18, 60 + 10 * w + k) for k in range(6)] for w in range(5)]
    pending, res = [], {}
    for wave in waves:
        ids = [(engine.submit(j), j) for j in wave]
        # take the previous wave's results while this wave's cohorts start
        for jid, j in pending:
            assert engine.wait(jid, 60.0) == 0
            res[jid] = (engine.get_result(jid, j), j)
            assert engine.release(jid) == 0
        pending = ids
    for jid, j in pending:
        assert engine.wait(jid, 60.0) == 0
        res[jid] = (engine.get_result(jid, j), j)
        assert engine.release(jid) == 0
    assert len(res) == 30
    for jid in list(res)[::5]:
        r, j = res[jid]
        alone = engine.scf_batch([j])[0]
        assert r["status"] == 0 and alone["status"] == 0
        assert abs(r["energy"] - alone["energy"]) * EV <= 1e-9
        assert r["iterations"] == alone["iterations"]


def test_batch_concurrent_with_served(engine):
    """nddo_scf_batch on the caller's thread while another thread submits,
    fetches and releases served requests whose cohorts start, capture their
    iteration graphs and retire: no call fails, every served result equals the
    alone solve, and every batch equals the first one bit for bit."""
    import threading
    print("batch concurrent with served", flush=True)
    served_jobs = [_job("water", 3 + (5 * k) % 17, 200 + k) if k % 2 else _job("alkane", 2 + (7 * k) % 19, 200 + k)
                   for k in range(24)]
    batch_jobs = [_job("alkane", 3 + k, 300 + k, precision="mixed") if k % 2 else _job("water", 4 + k, 300 + k)
                  for k in range(12)]
    out, errors = {}, []

    def client():
        try:
            for w in range(4):
                ids = [(engine.submit(j), j) for j in served_jobs[6 * w:6 * w + 6]]
                for jid, j in ids:
                    assert engine.wait(jid, 120.0) == 0
                    out[jid] = (engine.get_result(jid, j), j)
                    assert engine.release(jid) == 0
                time.sleep(0.01)
        except Exception as ex:   # reported by the main thread
            errors.append(repr(ex))

    th = threading.Thread(target=client)
    th.start()
    batches = []
    try:
        while th.is_alive() or len(batches) < 2:
            batches.append(engine.scf_batch(batch_jobs))
    finally:
        th.join()
    assert not errors, errors[0]
    assert len(out) == 24
    for b in batches[1:]:
        for r, r0 in zip(b, batches[0]):
            assert r["status"] == 0 and r["energy"] == r0["energy"] and r["iterations"] == r0["iterations"]
            assert np.array_equal(r["eps"], r0["eps"])
    for jid in list(out)[::3]:
        r, j = out[jid]
        alone = engine.scf_batch([j])[0]
        assert r["status"] == 0 and alone["status"] == 0
        assert abs(r["energy"] - alone["energy"]) * EV <= 1e-9
        assert r["iterations"] == alone["iterations"]


def test_served_statuses(engine):
    from molds_amd.nddo import Engine, Job
    odd = Job("pm3", [6, 1, 1, 1], np.eye(4, 3) * 2.0)
    jid = engine.submit(odd)
    assert engine.poll(jid) == -2          # ODD_ELECTRONS, known at submit
    assert engine.release(jid) == 0
    assert engine.poll(jid) == -7          # BAD_ARG: released
    small = Engine(0, hbm_budget_bytes=1 << 20)
    try:
        jid = small.submit(_job("water", 20, 41))
        assert small.wait(jid, 10.0) == -4  # ADMISSION_DENIED
    finally:
        small.close()
    # a queued or running request can be released before it finishes
    jid = engine.submit(_job("water", 30, 42))
    assert engine.release(jid) == 0
