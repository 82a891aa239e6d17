#!/usr/bin/env python3
"""Benchmark of the exmarg segment analysis on one MI355X.

Workload (the configuration the headline metric is quoted on): one two-detector segment, T = 16 s at
4096 Hz, 20-1024 Hz (n_time = 131072, 16065 bins), Gaussian noise colored by the analytic aLIGO fit with
one TaylorF2 injection (M_det = 30, q = 0.8, chi = 0.1, network SNR 12); a 2000-template bank on a
20 x 10 x 10 grid over M_det in [10, 80], q in [0.2, 1], chi in [-0.5, 0.5]; all 53 nodes of synthetic
float32 tables in the reference store layout (39.6 GB, HBM-resident); pe_intrinsic with the 'volume'
prior and the reference defaults n_loud = n_random = 1024.

A step is one pe_intrinsic call on the segment (matched filter, time selection, every table read, the
node and bank assembly).  value = segments/s over the timed steps.  After the timed steps the same line
reports the serving form on the same store ('serving'), the reference's own timed workloads beside its
published times ('reference_workloads', see published_workloads) and the CPU baseline.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# vector-ALU issue peak of one MI355X: 256 CUs x 4 SIMDs x 16 lanes (a wave64 instruction every 4 cycles)
# at the 2.4 GHz peak engine clock, in lane-instructions per second
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
TABLE_BYTES = 746_877_600        # one (87, 49, 200, 219) float32 table


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n-loud", type=int, default=1024)
    p.add_argument("--n-random", type=int, default=1024)
    p.add_argument("--config", type=int, choices=[1, 2], default=2,
                   help="plain workload: 2 (2000-template grid bank, 53 nodes; the headline) or 1 (64-template "
                        "chirp-mass grid, one node)")
    p.add_argument("--grid", default="20,10,10", help="bank grid (M, q, chi)")
    p.add_argument("--nodes", type=int, default=53)
    p.add_argument("--cpu-templates", type=int, default=512, help="templates in the CPU-baseline sample")
    p.add_argument("--cpu-nodes", type=int, default=16, help="nodes (the most used) in the CPU-baseline gather sample")
    p.add_argument("--check-templates", type=int, default=8, help="templates checked against the oracle")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle check")
    p.add_argument("--no-precision-form", action="store_true",
                   help="plain: skip the precision form's run on the same segment after the timed steps")
    p.add_argument("--gather-kernel", type=int, default=0, help="0 automatic, 1 reference-order kernel, 2-4 band kernel (piece, wave, scheduled form)")
    p.add_argument("--piece", type=int, default=0, help="walked times per piece of the band kernel (0 default)")
    p.add_argument("--tile", type=int, default=0, help="reads per band-kernel workgroup (0 default)")
    p.add_argument("--precision", choices=["f64", "f32"], default="f64",
                   help="table interpolation of the band kernel (index arithmetic and log-sum-exp stay float64)")
    p.add_argument("--filter-mib", type=int, default=0, help="matched-filter workspace per transform (MiB, 0 default)")
    p.add_argument("--filter-precision", choices=["f64", "f32"], default="f64",
                   help="matched-filter transform (weighting, modulus and phase stay float64)")
    p.add_argument("--ingest-nodes", type=int, default=4, help="ingest: full-size tables written and read back")
    p.add_argument("--mode", choices=["plain", "serve", "stream", "batch", "ingest", "tables", "published"],
                   default="plain",
                   help="plain: pe_intrinsic on one segment per step (the headline); serve: the request loop "
                        "under Poisson arrivals; stream: pe_redshift over 8 s segments with an estimated spectrum; "
                        "batch: 256 segments of uneven length in batched launches (float32 form, autotuned); "
                        "published: the reference's own timed workloads, beside its published times")
    p.add_argument("--published", default="speed,twelve,three",
                   help="the reference's timed workloads to run (speed, twelve, three; '' for none): the whole "
                        "line of --mode published, a 'reference_workloads' entry of the plain line")
    p.add_argument("--segments", type=int, default=64, help="stream: segments in the timed run")
    p.add_argument("--group", type=int, default=32, help="stream: segments sharing one spectrum estimate")
    p.add_argument("--rate", type=float, default=50.0, help="serve: mean arrival rate (requests/s)")
    p.add_argument("--requests", type=int, default=40, help="serve: requests in the timed run")
    p.add_argument("--serve-requests", type=int, default=24,
                   help="plain: requests of the serving workload run after the headline steps (0: none)")
    p.add_argument("--seed", type=int, default=7, help="serve: seed of the request stream")
    return p.parse_args()


def workload(args):
    """the segment and bank rows of the plain workload: config 2 (default) is a 2000-template grid bank over
    every node; config 1 the 64-template chirp-mass grid of the reference's CPU-runnable case (its one
    node is placed by plain_store)"""
    from exmarg_amd import synth
    seg, truth, snr = synth.simulate_at_snr(1, 20261203, (30.0, 0.8, 0.1), 12.0, 16.0, 20.0, 1024.0)
    if getattr(args, "config", 2) == 1:
        rows = synth.chirp_mass_grid_bank(30.0, 0.8, 0.1, 64)
    else:
        rows = synth.grid_bank(shape=tuple(int(v) for v in args.grid.split(",")))
    return seg, rows, float(snr[0])


def plain_store(args, bank):
    """the tables of the plain workload: every node (config 2), or config 1's one node at s = 4 max u, so
    that every template reaches it"""
    from exmarg_amd.search import store as stm
    if args.config == 1:
        return stm.synthetic_store(seed=0, s_nodes=np.array([4.0 * bank.u.max()]))
    return stm.synthetic_store(n_nodes=args.nodes, seed=0)


def measured_traffic(kernel_ms):
    """HBM bytes per launch of the gather kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    record (profiles/gather_traffic.json), used only when it was taken on this workload with the
    kernel source now in the tree"""
    import hashlib
    try:
        with open(os.path.join(ROOT, "profiles", "gather_traffic.json")) as f:
            rec = json.load(f)
        with open(os.path.join(ROOT, "exmarg_amd", "csrc", "k_band.hip"), "rb") as f:
            sha = hashlib.sha256(f.read()).hexdigest()[:16]
    except (OSError, ValueError):
        return None
    if rec.get("source_sha256_16") != sha or rec.get("workload") != "16s_2000tmpl_53nodes_walk2048":
        return None
    return rec


def segment_bytes(F, n_tmpl, n_time, n_nodes):
    """algorithmic bytes of one segment: data, template factors, filter output written and read once,
    every table read once"""
    return 32 * F + 8 * n_tmpl * F + 32 * n_tmpl * n_time + TABLE_BYTES * n_nodes


def cpu_leg(args, seg, bank, store, res, an_geometry, seed_row):
    """CPU baseline (oracle restatement, OpenMP) on a bounded sample, and the oracle check of a few
    templates' ln Lambda"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import exm_oracle as O
    from exmarg_amd.search import read_segments
    O.lib()
    rseg = read_segments(seg)
    d = {I: rseg["d_" + I][0] for I in "HL"}
    s2 = {I: rseg["sig2_" + I] for I in "HL"}
    nm, nt = bank.u.size, bank.n_time
    K, S = args.n_loud, args.n_random
    valid, lna, yN, bnode, blna = an_geometry
    nn = valid.shape[0]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    full_pick = O.draw_picks(nm, nt, K, S, seed_row)
    tabs_synth = O.Tables(synth=dict(seed=0, kind="smooth", nodes=np.arange(nn)), shape=store.table_shape(),
                          tau_idx=store.tau_idx)
    beta, delta = store.beta[0], store.delta[0]

    def jobs_for(sub, nodes=None):
        ids, tabs, lnas, rows = [], [], [], []
        for k in range(nn):
            if nodes is not None and k not in nodes:
                continue
            for j, m in enumerate(sub):
                if valid[k, m]:
                    ids.append(j); tabs.append(k); lnas.append(lna[k, m]); rows.append((k, j))
        for j, m in enumerate(sub):
            if bnode[m] >= 0 and (nodes is None or bnode[m] in nodes):
                ids.append(j); tabs.append(bnode[m]); lnas.append(blna[m]); rows.append((nn, j))
        return np.array(ids), np.array(tabs)[:, None], np.array(lnas)[:, None], rows

    # -- oracle check of a few templates (tables by the synthetic recipe, every node)
    chk = np.linspace(0, nm - 1, args.check_templates).astype(int)
    trc = O.filter_fp64(bank.W[chk], bank.lo, nt, d, s2)
    tmc = O.select_times(trc, store.tau_idx, K, S, 0, pick=full_pick[chk])
    ids, tabs, lnas, rows = jobs_for(chk)
    vals = O.gather(trc, tabs_synth, beta, delta, ids, lnas, tabs, tmc)[0]
    lamN = np.full((nn + 1, chk.size), -np.inf)
    for v, (k, j) in zip(vals, rows):
        lamN[k, j] = v
    lnW = O.bank_weights(yN, "volume", 0.1267)[:, chk]
    v = lnW + lamN
    mx = v.max(axis=0)
    lnl_ref = mx + np.log(np.exp(v - mx).sum(axis=0))
    err = float(np.max(np.abs(lnl_ref - res["ln_lambda"][0][chk])))

    # -- CPU baseline in the reference's own CPU arithmetic (src/exmarg/search/bank.py:66-75 complex64
    # numpy filter, single-threaded as numpy's FFT is; gather.py:396-420 float32 score ranked by
    # argpartition; gather.py:32-100 float64 gather with (template, lag) jobs over the threads as the
    # reference's prange): filter + selection of a template sample, gather rate on the most used nodes in RAM
    samp = np.linspace(0, nm - 1, args.cpu_templates).astype(int)
    t0 = time.perf_counter()
    trs = O.filter_ref(bank.W[samp], bank.lo, nt, d, s2)
    t_filter = time.perf_counter() - t0
    t0 = time.perf_counter()
    tms = O.select_times_ref(trs, store.tau_idx, K, S, 0)
    t_select = time.perf_counter() - t0
    counts = valid.sum(axis=1)
    nodes = list(np.argsort(-counts)[:max(1, args.cpu_nodes)])
    host_tables = []
    for k in range(nn):
        host_tables.append(O.synth_table(0, k, store.table_shape(), store.tau_idx) if k in nodes else None)
    ids, tabs, lnas, _ = jobs_for(samp, nodes=set(nodes))
    tabs_mem = O.Tables(tables=[t if t is not None else host_tables[nodes[0]] for t in host_tables],
                        tau_idx=store.tau_idx)
    t0 = time.perf_counter()
    O.gather(trs, tabs_mem, beta, delta, ids, lnas, tabs, tms, order="reference")
    t_gather = time.perf_counter() - t0
    lookups_sample = ids.size * store.tau_idx.size * (K + S)
    cpu_lookup_rate = lookups_sample / t_gather
    n_jobs_all = int(valid.sum() + (bnode >= 0).sum())
    lookups_all = n_jobs_all * store.tau_idx.size * (K + S)
    t_seg = (t_filter + t_select) * nm / samp.size + lookups_all / cpu_lookup_rate
    sample = ("the reference's CPU arithmetic (complex64 numpy filter, float32 score with argpartition, float64 "
              "OpenMP gather over (template, lag) jobs): filter+selection of %d of %d templates (scaled x%.1f) "
              "and the gather of those templates on the %d most-used nodes (%d reads x 87 lags x %d times, tables "
              "in RAM), scaled to all %d reads; %.1f s of CPU work timed" % (samp.size, nm, nm / samp.size, len(nodes),
                                                                           ids.size, K + S, n_jobs_all,
                                                                           t_filter + t_select + t_gather))
    # the same on one core: selection of a quarter of the sample, gather of it on one node
    O.set_threads(1)
    try:
        q = samp[: max(1, samp.size // 4)]
        trq = O.filter_ref(bank.W[q], bank.lo, nt, d, s2)
        t0 = time.perf_counter()
        tq = O.select_times_ref(trq, store.tau_idx, K, S, 0)
        t_sel1 = time.perf_counter() - t0
        ids1, tabs1, lnas1, _ = jobs_for(q, nodes={nodes[0]})
        t0 = time.perf_counter()
        O.gather(trq, tabs_mem, beta, delta, ids1, lnas1, tabs1, tq, order="reference")
        rate1 = ids1.size * store.tau_idx.size * (K + S) / (time.perf_counter() - t0)
    finally:
        O.set_threads(threads)
    t_seg1 = t_filter * nm / samp.size + t_sel1 * nm / q.size + lookups_all / rate1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    cpu = dict(value=1.0 / t_seg, unit="segments/s", cores=threads, kind="port", sample=sample,
               seconds_per_segment=t_seg, gather_lookups_per_s=cpu_lookup_rate, cpu_model=model,
               host_cpus=os.cpu_count(),
               one_core=dict(value=1.0 / t_seg1, seconds_per_segment=t_seg1, gather_lookups_per_s=rate1,
                             sample="selection of %d templates and their gather on 1 node (%d reads), 1 thread; "
                                    "the filter (numpy, single-threaded) as above" % (q.size, ids1.size)))
    return cpu, err, chk.size


def serve_requests(n, seed, rate):
    """the serving workload: Poisson arrivals at `rate`; T uniform over {4, 8, 16, 32, 64} s at 20-1024 Hz;
    banks of n_tmpl log-uniform in [500, 8000] on (n_tmpl / 20) x 5 x 4 grids over M_det in [10, 80],
    q in [0.2, 1], chi in [-0.5, 0.5]; 30 % with one injection of network SNR log-uniform in [6, 40]
    drawn inside the bank range, the rest noise only"""
    from exmarg_amd import synth
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        T = float(rng.choice([4.0, 8.0, 16.0, 32.0, 64.0]))
        n_tmpl = int(np.exp(rng.uniform(np.log(500.0), np.log(8000.0))))
        nM = max(2, int(round(n_tmpl / 20.0)))
        rows = synth.grid_bank(shape=(nM, 5, 4))
        if rng.uniform() < 0.3:
            params = (float(rng.uniform(12.0, 75.0)), float(rng.uniform(0.25, 1.0)), float(rng.uniform(-0.45, 0.45)))
            snr = float(np.exp(rng.uniform(np.log(6.0), np.log(40.0))))
            seg, _, _ = synth.simulate_at_snr(1, seed * 1000 + i, params, snr, T, 20.0, 1024.0)
        else:
            seg, _, _ = synth.simulate(1, seed * 1000 + i, lambda r: (30.0, 0.8, 0.1), lambda r: 0.0, T, 20.0, 1024.0)
        out.append((float(t[i]), seg, rows))
    return out


def serve_run(st, n_requests, rate, seed, n_loud=1024, n_random=1024):
    """the request loop on the resident store under Poisson arrivals (serve_requests): throughput and
    latency of the requests, after one warm-up request per segment length"""
    from exmarg_amd import _hip
    from exmarg_amd.search import Server
    reqs = serve_requests(n_requests, seed, rate)
    srv = Server(st, n_loud=n_loud, n_random=n_random, max_request=(8000, 8 * 65536, 64257))
    # warm-up: one small request per segment length (FFT plans, kernel loads)
    for T in (4.0, 8.0, 16.0, 32.0, 64.0):
        for t_arr, seg, rows in reqs:
            if seg["T"] == T:
                srv.submit(seg, rows[:40 * 20][::10], seed=1)
                break
    srv.run()
    srv.stats.update(batches=0, admitted=0, admitted_during_gather=0, finished=0, refused=0)
    _hip.profile(True)
    t0 = time.perf_counter()
    pending = list(reqs)
    handles = []
    while pending or not srv.idle:
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            t_arr, seg, rows = pending.pop(0)
            handles.append(srv.submit(seg, rows, seed=len(handles), t_submit=t0 + t_arr))
        if not srv.step():
            time.sleep(2e-4)
    wall = time.perf_counter() - t0
    g_ms, g_n = _hip.profile_get("gather")
    f_ms, f_n = _hip.profile_get("filter")
    _hip.profile(False)
    srv.close()
    ok = [h for h in handles if h.error is None]
    lat = np.array([h.latency for h in ok])
    span = max(h.t_done for h in ok) - t0 - reqs[0][0]
    n_tmpl = np.array([h.rows.shape[0] for h in ok])
    return {"value": len(ok) / span, "unit": "requests/s", "offered_rate": rate, "requests": len(handles),
            "finished": len(ok), "refused": len(handles) - len(ok),
            "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                          "max": float(lat.max())},
            "wall_s": wall, "batches": srv.stats["batches"],
            "admitted_during_gather": srv.stats["admitted_during_gather"],
            "templates_mean": float(n_tmpl.mean()), "gather_ms_total": g_ms, "filter_ms_total": f_ms,
            "gpu_gather_busy_frac": g_ms / 1e3 / wall}


SERVE_WORKLOAD = ("Poisson arrivals; T in {4,8,16,32,64} s at 20-1024 Hz; banks of 500-8000 templates (log-uniform); "
                  "30%% injections SNR 6-40; 53 HBM-resident table nodes; pe_intrinsic 'volume', n_loud=n_random=%d")


def main_serve(args):
    """the request loop under Poisson arrivals: one JSON line with throughput and latency"""
    from exmarg_amd import _hip
    from exmarg_amd.search import store as stm
    _hip.init(0)
    info = _hip.device_info()
    # the float32 form on request (--precision f32: float32 transform and interpolation)
    prec = 1 if args.precision == "f32" else 0
    _hip.call("exm_set_gather_precision", prec)
    _hip.call("exm_set_filter_precision", prec)
    st = stm.synthetic_store(n_nodes=args.nodes, seed=0)
    res = serve_run(st, args.requests, args.rate, args.seed, args.n_loud, args.n_random)
    out = {"metric": "serving: requests/s and p50/p99 latency under Poisson arrivals", "n_gpus": 1,
           "higher_is_better": True}
    out.update(res)
    out.update({"dtype": ("f32 transform and interpolation, f64 weighting, index arithmetic and log-sum-exp" if prec
                          else "f64"), "data": "synthetic",
                "config": {"workload": SERVE_WORKLOAD % args.n_loud, "device": info["name"],
                           "num_cu": info["num_cu"]}})
    print(json.dumps(out), flush=True)


def stream_groups(n, group, seed, n_blocks=32, frac_inj=0.05, mdet_range=(10.0, 500.0)):
    """the stream workload: 8 s segments at 20-512 Hz in Gaussian noise (aLIGO fit), frac_inj of them with
    one equal-mass, zero-spin injection drawn from the population (GWTC-5 masses, Madau-Dickinson
    redshifts) with M_det on mdet_range at y = D0 / D_L(z); every group of `group` segments analysed with
    the mean of n_blocks periodograms times N/(N-2) as its noise variance"""
    from exmarg_amd import cosmology, synth
    from exmarg_amd.search import stream
    rng = np.random.default_rng(seed)
    groups, truths = [], []
    for g0 in range(0, n, group):
        k = min(group, n - g0)
        params = []
        for _ in range(k):
            if rng.uniform() < frac_inj:
                m1, z = stream.draw_population(rng, mdet_range)
                params.append((2.0 * m1 * (1.0 + z), m1, z, synth.D0 / float(np.ravel(cosmology.luminosity_distance(z))[0])))
            else:
                params.append(None)
        it_d, it_y = iter(params), iter(params)
        seg, _, snr = synth.simulate(k, seed * 100000 + g0, lambda r: (next(it_d) or (30.0,))[0:1] + (1.0, 0.0),
                                     lambda r: (lambda p: 0.0 if p is None else p[3])(next(it_y)), 8.0, 20.0, 512.0)
        seg["sig2"] = stream.psd_estimate(stream.noise_periodograms(seg["sig2"], n_blocks, rng))
        groups.append(seg)
        truths.append([(p, float(s)) for p, s in zip(params, snr)])
    return groups, truths


def stream_strain(groups, fs=4096.0, n_blocks=32, seed=0):
    """the stream workload as strain: each group's segments as consecutive blocks of band-limited strain
    sampled at fs (x = irfft of d fs on the band bins, so that the front end gives d back), and n_blocks
    neighbouring noise-only blocks per detector drawn with the true noise variance (aLIGO fit), from
    which the front end estimates the group's noise variance"""
    from exmarg_amd import synth
    rng = np.random.default_rng(seed)
    out = []
    for seg in groups:
        T = float(seg["T"])
        n = int(round(T * fs))
        nu = np.asarray(seg["nu"])
        sig2 = T * synth.psd_ajith(nu / T) / 2.0
        strain, noise = {}, {}
        for det in "HL":
            Y = np.zeros((seg["d_" + det].shape[0], n // 2 + 1), complex)
            Y[:, nu] = seg["d_" + det] * fs
            strain[det] = np.fft.irfft(Y, n, axis=1).ravel()
            Y = np.zeros((n_blocks, n // 2 + 1), complex)
            Y[:, nu] = fs * np.sqrt(sig2 / 2.0) * (rng.standard_normal((n_blocks, nu.size))
                                                   + 1j * rng.standard_normal((n_blocks, nu.size)))
            noise[det] = np.fft.irfft(Y, n, axis=1).ravel()
        out.append((strain, noise, T, float(seg["fmin"]), float(seg["fmax"])))
    return out


def main_stream(args):
    """pe_redshift over a stream of 8 s segments given as strain: one JSON line with segments/s"""
    from exmarg_amd import _hip
    from exmarg_amd.priors import madau_dickinson_ymin
    from exmarg_amd.search import pe_redshift_stream, stream, store as stm
    _hip.init(0)
    info = _hip.device_info()
    prec = 1 if args.precision == "f32" else 0            # the float32 form on request
    _hip.call("exm_set_gather_precision", prec)
    _hip.call("exm_set_filter_precision", prec)
    st = stm.synthetic_store(n_nodes=args.nodes, seed=0)
    masses = np.geomspace(10.0, 500.0, 3517)
    fs = 4096.0
    groups, truths = stream_groups(args.segments + args.group, args.group, args.seed)
    strains = stream_strain(groups, fs)
    warm, strains, truths = strains[:1], strains[1:], truths[1:]
    kw = dict(zmax=10.0, nz=1600, m1_range=(2.0, 200.0), nm1=1600, mdet_range=(10.0, 500.0),
              ymin=madau_dickinson_ymin(), n_loud=args.n_loud, n_random=args.n_random)
    front_s = [0.0]

    def front(items):
        """the segments and the noise estimate of each group from its strain, on the device"""
        for x, noise, T, fmin, fmax in items:
            t1 = time.perf_counter()
            g = stream.segments_from_strain(x, noise, fs, T, fmin, fmax)
            front_s[0] += time.perf_counter() - t1
            yield g

    for _ in pe_redshift_stream(front(warm), masses, st, **kw):
        pass
    front_s[0] = 0.0
    inj = {(g, j) for g, tr in enumerate(truths) for j, (p, _) in enumerate(tr) if p is not None}
    _hip.profile(True)
    t0 = time.perf_counter()
    res = list(pe_redshift_stream(front(strains), masses, st, keep=lambda g, row: (g, row) in inj, **kw))
    wall = time.perf_counter() - t0
    g_ms, g_n = _hip.profile_get("gather")
    sf_ms, sf_n = _hip.profile_get("source_frame")
    _hip.profile(False)
    nseg = sum(r["ln_b"].size for r in res)
    ln_b = np.concatenate([r["ln_b"] for r in res])
    m1, z = res[0]["m1"], res[0]["z"]
    prior_mz = stream.population_prior(m1, z, (10.0, 500.0))
    levels = []
    for g, r in enumerate(res):
        for row, L in r["kept"].items():
            p, snr = truths[g][row]
            levels.append(stream.posterior_levels(L, prior_mz, m1, z, p[1], p[2]) + (snr,))
    out = {
        "metric": "stream: segments/s of pe_redshift over 8 s segments with an estimated noise spectrum",
        "value": nseg / wall, "unit": "segments/s", "n_gpus": 1, "higher_is_better": True,
        "segments": nseg, "wall_s": wall, "ms_per_segment": 1e3 * wall / nseg,
        "gather_ms_per_segment": g_ms / nseg, "source_frame_ms_per_segment": sf_ms / max(sf_n, 1),
        "front_end_ms_per_segment": 1e3 * front_s[0] / nseg,
        "ln_b": {"median": float(np.median(ln_b)), "max": float(ln_b.max())},
        "injections": [{"m1_level": a, "z_level": b, "snr": c} for a, b, c in levels],
        "dtype": ("f32 transform and interpolation, f64 weighting, index arithmetic and log-sum-exp" if prec
                  else "f64"), "data": "synthetic",
        "config": {"workload": "8 s segments at 20-512 Hz given as strain sampled at 4096 Hz, 5%% population "
                               "injections (equal mass, zero spin); segments and noise variance (mean of the "
                               "periodograms of 32 neighbouring 8 s noise blocks x 32/30, per group of %d) formed "
                               "on the device from the strain; 3517-mass bank "
                               "10-500 Msun (geometric); 53 HBM-resident table nodes; pe_redshift on a 1600 x 1600 "
                               "(m1, z) grid, z <= 10, ymin %.4f, n_loud=n_random=%d"
                               % (args.group, madau_dickinson_ymin(), args.n_loud),
                   "device": info["name"], "num_cu": info["num_cu"]},
    }
    print(json.dumps(out), flush=True)


def batch_items(n, seed):
    """the batched workload: n segments with T uniform over {4, 8, 16, 32, 64} s at 20-1024 Hz, half noise
    only, half with one TaylorF2 injection (M_det, q, chi inside the bank, network SNR log-uniform in
    [8, 30]); one 2000-template 20 x 10 x 10 bank per segment length"""
    from exmarg_amd import synth
    rng = np.random.default_rng(seed)
    Ts = rng.choice([4.0, 8.0, 16.0, 32.0, 64.0], n)
    items = []
    for T in (4.0, 8.0, 16.0, 32.0, 64.0):
        k = int((Ts == T).sum())
        if k == 0:
            continue
        params = [(float(rng.uniform(12.0, 75.0)), float(rng.uniform(0.25, 1.0)), float(rng.uniform(-0.45, 0.45)))
                  for _ in range(k)]
        snr = [float(np.exp(rng.uniform(np.log(8.0), np.log(30.0)))) if rng.uniform() < 0.5 else 0.0
               for _ in range(k)]
        it_p = iter(params)
        seg1, _, snr1 = synth.simulate(k, seed * 1000 + int(T), lambda r: next(it_p), lambda r: 1.0, T, 20.0, 1024.0)
        it_p, it_y = iter(params), iter(np.array(snr) / snr1)
        seg, _, _ = synth.simulate(k, seed * 1000 + int(T), lambda r: next(it_p), lambda r: float(next(it_y)), T,
                                   20.0, 1024.0)
        items.append((seg, T))
    return items


def main_batch(args):
    """pe_intrinsic_batch over uneven segments: one JSON line with segments/s, the plain drop-in's rate on
    the same segments and the float32 form's ln B error against the float64 form"""
    from exmarg_amd import _hip, tune
    from exmarg_amd.search import SegmentBatch, mass_bank, pe_intrinsic, pe_intrinsic_batch, store as stm
    from exmarg_amd.search.batch import release_kept
    from exmarg_amd.search import pe as _pe
    from exmarg_amd.search.segments import read_segments
    from exmarg_amd import synth
    _hip.init(0)
    info = _hip.device_info()
    st = stm.synthetic_store(n_nodes=args.nodes, seed=0)
    raw = batch_items(args.segments, args.seed)
    rows = synth.grid_bank(shape=(20, 10, 10))
    items = [(seg, mass_bank(rows, seg, "hip")) for seg, T in raw]
    kw = dict(n_loud=args.n_loud, n_random=args.n_random, seed=0)
    kwb = dict(kw, keep_buffer=True)        # a batched pipeline keeps its batch buffer from call to call
    # the plain drop-in (float64, default kernel form, one segment per call) on a sample of the same
    # segments, run first: the batch slots below take most of the HBM, and the first allocations after
    # their release wait for the driver to clear that memory.  The first call of each bank (its node
    # geometry, weights and read plan, the work memory grown to its length) is timed apart from the
    # calls after it.
    plain_n, t_plain, t_first = 0, 0.0, []
    for seg, bank in items:
        nrow = read_segments(seg)["rows"].size
        t1 = time.perf_counter()
        pe_intrinsic(seg, bank, st, rows=[0], resident=True, **kw)
        t_first.append(time.perf_counter() - t1)
        for j in range(min(3, nrow)):
            t1 = time.perf_counter()
            pe_intrinsic(seg, bank, st, rows=[j], resident=True, **kw)
            t_plain += time.perf_counter() - t1
            plain_n += 1
    plain = plain_n / t_plain
    plain_cold = len(t_first) / sum(t_first)
    prec = 1 if args.precision == "f32" else 0
    _hip.call("exm_set_gather_precision", prec)
    _hip.call("exm_set_filter_precision", prec)
    # autotune the band kernel on a sample batch (8 segments of the 16 s bank, or the first bank)
    sample_seg, sample_bank = next(((s, b) for s, b in items if s["T"] == 16.0), items[0])
    rseg = read_segments(sample_seg)
    an = _pe._Analysis(sample_bank, st, 0.1267, True)
    an.start()
    sb = SegmentBatch(st, args.n_loud, args.n_random, max_templates=8 * sample_bank.u.size)
    for j in range(min(8, rseg["rows"].size)):
        sb.add(an, {I: rseg["d_" + I][j] for I in "HL"}, {I: rseg["sig2_" + I] for I in "HL"}, j)
    tuned = tune.tune_batch(sb, repeats=1)
    sb.free()
    # warm-up (one untimed run of the same workload: the work memory of its longest segments and the batch
    # buffer are allocated there, as a batched pipeline allocates them once), then the timed batched run
    pe_intrinsic_batch(items, st, **kwb)
    _hip.profile(True)
    t0 = time.perf_counter()
    res = pe_intrinsic_batch(items, st, **kwb)
    wall = time.perf_counter() - t0
    g_ms, g_n = _hip.profile_get("gather")
    f_ms, _ = _hip.profile_get("filter")
    _hip.profile(False)
    nseg = sum(r["ln_b"].size for r in res)
    # the float64 form of the same batch (the reference for the float32 error)
    _hip.call("exm_set_gather_precision", 0)
    _hip.call("exm_set_filter_precision", 0)
    res64 = pe_intrinsic_batch(items, st, **kwb)
    release_kept(st)
    dlnb = float(max(np.max(np.abs(a["ln_b"] - b["ln_b"])) for a, b in zip(res, res64)))
    dlnl = float(max(np.nanmax(np.abs(a["ln_lambda"] - b["ln_lambda"])) for a, b in zip(res, res64)))
    out = {
        "metric": "batched: segments/s over uneven segments (float32 gather, float64 log-sum-exp)",
        "value": nseg / wall, "unit": "segments/s", "n_gpus": 1, "higher_is_better": True,
        "segments": nseg, "wall_s": wall, "gather_launches": g_n, "gather_ms_total": g_ms, "filter_ms_total": f_ms,
        "plain_dropin_segments_per_s": plain, "plain_sample_segments": plain_n, "speedup_vs_plain": nseg / wall / plain,
        "plain_first_call_segments_per_s": plain_cold, "speedup_vs_plain_first_call": nseg / wall / plain_cold,
        "float32_vs_float64": {"max_abs_dln_b": dlnb, "max_abs_dln_lambda": dlnl},
        "autotune": tuned,
        "dtype": "f32 transform and interpolation, f64 weighting, index arithmetic and log-sum-exp" if prec else "f64",
        "data": "synthetic",
        "config": {"workload": "%d segments, T uniform over {4,8,16,32,64} s at 20-1024 Hz, half noise, half one "
                               "TaylorF2 injection (SNR 8-30); 2000-template 20x10x10 bank per length; 53 "
                               "HBM-resident table nodes; pe_intrinsic 'volume', n_loud=n_random=%d"
                               % (nseg, args.n_loud),
                   "device": info["name"], "num_cu": info["num_cu"]},
    }
    print(json.dumps(out), flush=True)


def main_ingest(args):
    """store ingest: full-size tables written as bd_*.npz files, then loaded into HBM through host arrays
    (the reference's way) and straight from the files; one JSON line with both rates"""
    import shutil
    import tempfile
    from exmarg_amd import _hip, synth
    from exmarg_amd.search import store as stm
    _hip.init(0)
    info = _hip.device_info()
    n = args.ingest_nodes
    tmp = tempfile.mkdtemp(prefix="exm_store_")
    try:
        src = stm.synthetic_store(n_nodes=n, seed=0)
        beta, delta, _ = synth.store_grids()
        arrays = dict(s=src.s, beta=beta, delta=delta, tau_idx=src.tau_idx, alpha=synth.ALPHA.copy(),
                      tables=[src.device.read_table(k) for k in range(n)])
        stm.free(src)
        synth.write_store(tmp, arrays)
        nbytes = sum(T.nbytes for T in arrays["tables"])
        del arrays
        t0 = time.perf_counter()
        a = stm.load_redshift_store(tmp, readers=6)
        t1 = time.perf_counter()
        stm.upload(a)
        t2 = time.perf_counter()
        stm.free(a)
        del a
        t3 = time.perf_counter()
        b = stm.load_redshift_store(tmp, readers=6, host=False)
        stm.upload(b)
        t4 = time.perf_counter()
        stm.free(b)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out = {"metric": "store ingest: GB/s from bd_*.npz files to HBM-resident device layout",
           "value": nbytes / (t4 - t3) / 1e9, "unit": "GB/s", "n_gpus": 1, "higher_is_better": True,
           "tables": n, "bytes": nbytes,
           "direct_s": t4 - t3, "host_path_s": {"read": t1 - t0, "upload": t2 - t1, "total": t2 - t0},
           "host_path_GB_per_s": nbytes / (t2 - t0) / 1e9,
           "note": "files freshly written, so reads come from the page cache; 6 reader threads both ways",
           "data": "synthetic", "config": {"workload": "%d full (87, 49, 200, 219) float32 tables" % n,
                                           "device": info["name"]}}
    print(json.dumps(out), flush=True)


def main_tables(args):
    """table build, stage 3 (the sum over the orders): one node at the reference geometry (121 orders, 87 lags,
    110 q points, 49 phases -> 87 x 49 x 200 x 219 float32) from synthetic stage-2 slices; nodes/s on the GPU
    with the host copies, the kernel time, and the oracle on a bounded sample of lags"""
    import math
    from exmarg_amd import _hip, build_bd
    _hip.init(0)
    info = _hip.device_info()
    nord, ntau, nq, nphif = 121, 87, 110, 49
    lnq = np.log(np.geomspace(100.0 / 2000.0, 100.0, nq))
    n = np.arange(nord, dtype=np.float64)[:, None, None, None]
    lgam = np.array([math.lgamma(k + 1.0) for k in range(nord)])[:, None, None, None]
    cph = np.cos(np.arange(nphif) * np.pi / 48.0)[None, None, None, :]
    lag = (np.arange(ntau) / ntau)[None, :, None, None]
    col = (2.0 * n * lnq[-1] - lgam + 0.3 * n * cph - lag - 0.02 * n * np.arange(nq)[None, None, :, None]).astype(np.float32)
    row = (2.0 * n * lnq[None, None, :, None] - lgam + 0.3 * n * cph - lag).astype(np.float32)
    ladder = [nord - 1]
    build_bd.sum_over_orders(col, row, lnq, ladder)                     # warm-up
    _hip.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = build_bd.sum_over_orders(col, row, lnq, ladder)
    wall = (time.perf_counter() - t0) / args.steps
    k_ms, k_n = _hip.profile_get("tables")
    _hip.profile(False)
    kernel_ms = k_ms / max(k_n, 1)
    beta, delta = build_bd.grids(lnq)
    outputs = ntau * nphif * beta.size * delta.size
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
        import exm_oracle
        lags = 2
        c0 = time.perf_counter()
        ref = exm_oracle.bd_sum_orders(col[:, :lags], row[:, :lags], lnq, ladder)
        cs = (time.perf_counter() - c0) * ntau / lags
        same = float(np.mean(ref[ladder[0]] == out[ladder[0]][:lags]))
        cpu = {"value": 1.0 / cs, "unit": "nodes/s", "cores": 1, "kind": "port",
               "sample": "the oracle (numpy, float64) on %d of the %d lags, scaled" % (lags, ntau),
               "seconds_per_node": cs, "identical_fraction": same}
    in_bytes = 2 * col.nbytes
    out_bytes = 4 * outputs * len(ladder)
    line = {"metric": "table build stage 3: nodes/s (sum over 121 orders -> one 87 x 49 x 200 x 219 float32 table)",
            "value": 1.0 / wall, "unit": "nodes/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
            "seconds_per_node": wall, "kernel_ms": kernel_ms,
            "exp_per_s": outputs * nord / (kernel_ms * 1e-3),
            "roofline": {"bound": "hbm", "achieved": (in_bytes + out_bytes) / (kernel_ms * 1e-3) / 1e9,
                         "peak": 8000.0, "unit": "GB/s",
                         "frac": (in_bytes + out_bytes) / (kernel_ms * 1e-3) / 8e12, "traffic": None,
                         "kernel": "k_bd_sum_orders", "bytes_per_launch": in_bytes + out_bytes},
            "cpu_baseline": cpu, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "one node: 121 orders, 87 lags, 110 q points, 49 phases -> 200 beta x 219 Delta",
                       "device": info["name"]}}
    print(json.dumps(line), flush=True)


# The reference's own timings of the workloads below (its notebooks, NVIDIA hardware, its real tables of the
# same geometry): seconds per segment, the GPU, where.
PUBLISHED = {
    "speed": (0.102, "NVIDIA L40S", "notebooks/05_speed.ipynb cell 7: 128 segments in 13 s, batches of 16"),
    "twelve": (2.75, "NVIDIA L40S", "notebooks/05_speed.ipynb cell 3: 12 segments in 33.0 s (chunk=12, "
                                    "block=64); 1.7 s per segment with the default block in "
                                    "02_single_segments.ipynb cell 14"),
    "three": (33.0, "GPU not named", "notebooks/02_single_segments.ipynb cell 5: 2 segments x 78,000 templates "
                                     "in 66 s (29.9 s per segment over 8 segments, cell 11)"),
}


def reference_noise():
    """S(f) of the noise curve in the reference's own segment files (sig2 = T S / 2 on their 8 s, 20-512 Hz
    grid; tests/golden/segments_2p.npz holds two of those segments)"""
    with np.load(os.path.join(ROOT, "tests", "golden", "segments_2p.npz")) as z:
        f, S = z["f"].copy(), 2.0 * z["sig2"] / float(z["T"])
    return lambda x: np.interp(x, f, S, left=np.inf, right=np.inf)


def reads_of(bank, st, ymin=0.1267):
    """(template, node) reads of one segment: the work of its gather, 87 lags x 2048 walked times each"""
    from exmarg_amd.search.pe import _node_geometry
    geometry = _node_geometry(bank, st, ymin)
    return int(geometry[0].sum() + (geometry[3] >= 0).sum())


def published_entry(name, seconds, n_seg, reads, n_lag, extra):
    ref_s, gpu, where = PUBLISHED[name]
    s = seconds / n_seg
    return dict(extra, seconds_per_segment=s, segments_per_gpu_hour=3600.0 / s, segments=n_seg,
                reads_per_segment=reads, lookups_per_segment=reads * n_lag * 2048,
                reference={"seconds_per_segment": ref_s, "gpu": gpu, "source": where},
                speedup_vs_reference=ref_s / s)


def published_workloads(st, which):
    """the reference's own timed workloads through the drop-in on the resident store st, beside its published
    times (PUBLISHED):

    speed   128 simulated 8 s segments (20-512 Hz, equal-mass injections M_det ~ U(20, 200), D_L log-uniform in
            1-15 Gpc), the 120 masses geomspace(20, 200) of its paper, pe_redshift on the 1600 x 1600 grid in
            batches of 16 segments sorted by SNR (05_speed.ipynb cell 7)
    twelve  12 segments of 8 s with its 3517-mass bank, pe_redshift in one call (05_speed.ipynb cell 3): the two
            segments of its data file in tests/golden, six times each
    three   2 segments of 4 s (20-512 Hz) with its 78,000-template bank (120 masses geomspace(20, 200) x 26 q in
            [0.3, 1] x 25 chi in [-0.6, 0.6]) in 10 pieces, pe_intrinsic 'volume' (02_single_segments.ipynb cell 5)

    The segments' noise is the curve of the reference's segment files; the tables are the synthetic 53 nodes
    of the same geometry (the work of a lookup does not depend on the table values)."""
    from exmarg_amd import synth
    from exmarg_amd.search import mass_bank, pe_intrinsic, pe_redshift
    noise = reference_noise()
    n_lag = st.tau_idx.size
    out = {}
    fast = dict(block=64, resident=True)
    if "speed" in which:
        draw = lambda rng: (rng.uniform(20.0, 200.0), 1.0, 0.0)                       # noqa: E731
        draw_y = lambda rng: synth.D0 / np.exp(rng.uniform(np.log(1000.0), np.log(15000.0)))   # noqa: E731
        seg, _, snr = synth.simulate(128, 20261020, draw, draw_y, 8.0, 20.0, 512.0, psd=noise)
        bank = mass_bank(np.geomspace(20.0, 200.0, 120), seg, "hip")
        order = np.argsort(snr)
        batches = [np.sort(order[i:i + 16]) for i in range(0, order.size, 16)]
        pe_redshift(seg, bank, st, rows=batches[0], chunk=16, **fast)      # warm-up: kernels and buffers
        per = []
        t0 = time.perf_counter()
        for b in batches:
            t1 = time.perf_counter()
            pe_redshift(seg, bank, st, rows=b, chunk=16, **fast)
            per.append((time.perf_counter() - t1) / b.size)
        wall = time.perf_counter() - t0
        out["speed"] = published_entry(
            "speed", wall, seg["d_H"].shape[0], reads_of(bank, st), n_lag,
            {"workload": "128 simulated 8 s segments at 20-512 Hz, SNR %.1f-%.1f; 120-mass equal-mass bank "
                         "geomspace(20, 200); pe_redshift 1600 x 1600, batches of 16" % (snr.min(), snr.max()),
             "batch_seconds_per_segment": {"min": min(per), "max": max(per)}})
        print("speed: %.4f s per segment" % (wall / 128), file=sys.stderr, flush=True)
    if "twelve" in which:
        with np.load(os.path.join(ROOT, "tests", "golden", "segments_2p.npz")) as z:
            two = {k: z[k] for k in z.files}
        seg = dict(two, d_H=np.tile(two["d_H"], (6, 1)), d_L=np.tile(two["d_L"], (6, 1)))
        bank = mass_bank(os.path.join(ROOT, "exmarg_amd", "data", "bank_2p.npy"), seg, "hip")
        pe_redshift(seg, bank, st, rows=[0, 1], chunk=12, **fast)           # warm-up: kernels and buffers
        t0 = time.perf_counter()
        res = pe_redshift(seg, bank, st, chunk=12, **fast)
        wall = time.perf_counter() - t0
        out["twelve"] = published_entry(
            "twelve", wall, 12, reads_of(bank, st), n_lag,
            {"workload": "12 segments of 8 s at 20-512 Hz (the reference's two_parameters rows 0 and 1, six times "
                         "each); its 3517-mass bank; pe_redshift 1600 x 1600 in one call",
             "ln_b": [float(v) for v in res["ln_b"][:2]]})
        print("twelve: %.4f s per segment" % (wall / 12), file=sys.stderr, flush=True)
    if "three" in which:
        draw = lambda rng: (rng.uniform(25.0, 150.0), rng.uniform(0.35, 1.0), rng.uniform(-0.5, 0.5))  # noqa: E731
        draw_y = lambda rng: synth.D0 / np.exp(rng.uniform(np.log(400.0), np.log(2000.0)))   # noqa: E731
        seg, _, snr = synth.simulate(2, 20261021, draw, draw_y, 4.0, 20.0, 512.0, psd=noise)
        axes = np.meshgrid(np.geomspace(20.0, 200.0, 120), np.linspace(0.3, 1.0, 26), np.linspace(-0.6, 0.6, 25),
                           indexing="ij")
        rows = np.stack([a.ravel() for a in axes], axis=1)
        pieces = np.array_split(np.arange(rows.shape[0]), 10)
        t0 = time.perf_counter()
        banks = [mass_bank(rows[p], seg, "hip") for p in pieces]
        t_bank = time.perf_counter() - t0
        kw = dict(prior="volume", resident=True)
        pe_intrinsic(seg, banks[0], st, rows=[0, 1], chunk=2, **kw)         # warm-up: kernels and buffers
        lnl = np.empty((2, rows.shape[0]))
        t0 = time.perf_counter()
        for p, b in zip(pieces, banks):
            lnl[:, p] = pe_intrinsic(seg, b, st, rows=[0, 1], chunk=2, **kw)["ln_lambda"]
        wall = time.perf_counter() - t0
        out["three"] = published_entry(
            "three", wall, 2, sum(reads_of(b, st) for b in banks), n_lag,
            {"workload": "2 simulated 4 s segments at 20-512 Hz (SNR %.1f, %.1f); 78,000-template (M_det, q, chi) "
                         "bank in 10 pieces; pe_intrinsic 'volume'" % tuple(snr),
             "bank_build_s": t_bank})
        print("three: %.3f s per segment" % (wall / 2), file=sys.stderr, flush=True)
    return out


def main_published(args):
    """the reference's own timed workloads (published_workloads): one JSON line"""
    from exmarg_amd import _hip
    from exmarg_amd.search import store as stm
    _hip.init(0)
    info = _hip.device_info()
    st = stm.synthetic_store(n_nodes=args.nodes, seed=0)
    which = [w.strip() for w in args.published.split(",") if w.strip()]
    out = {"metric": "seconds per segment of the reference's own timed workloads (its notebooks), this drop-in "
                     "on one MI355X beside the reference's published times",
           "unit": "s/segment", "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
           "device": info["name"], "num_cu": info["num_cu"], "workloads": published_workloads(st, which)}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the path runs one process per GPU and this build targets a single "
              "MI355X (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.mode == "serve":
        return main_serve(args)
    if args.mode == "stream":
        return main_stream(args)
    if args.mode == "ingest":
        return main_ingest(args)
    if args.mode == "tables":
        return main_tables(args)
    if args.mode == "published":
        return main_published(args)
    if args.mode == "batch":
        if args.precision == "f64" and "--precision" not in " ".join(sys.argv):
            args.precision = "f32"
        if "--segments" not in " ".join(sys.argv):
            args.segments = 256
        return main_batch(args)
    from exmarg_amd import _hip
    from exmarg_amd.search import mass_bank, pe_intrinsic, store as stm
    from exmarg_amd.search.pe import _node_geometry

    seg, rows, snr = workload(args)
    _hip.init(0)
    _hip.call("exm_set_gather_kernel", args.gather_kernel)
    _hip.call("exm_set_gather_piece", args.piece)
    _hip.call("exm_set_gather_tile", args.tile)
    _hip.call("exm_set_gather_precision", 1 if args.precision == "f32" else 0)
    _hip.call("exm_set_filter_workspace", args.filter_mib << 20)
    _hip.call("exm_set_filter_precision", 1 if args.filter_precision == "f32" else 0)
    info = _hip.device_info()
    bank = mass_bank(rows, seg, "hip")
    st = plain_store(args, bank)
    kw = dict(prior="volume", ymin=0.1267, n_loud=args.n_loud, n_random=args.n_random, seed=0, resident=True)
    for _ in range(args.warmup):
        pe_intrinsic(seg, bank, st, **kw)
    _hip.profile(True)
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        res = pe_intrinsic(seg, bank, st, **kw)
        lat.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t_start
    g_ms, g_n = _hip.profile_get("gather")
    f_ms, f_n = _hip.profile_get("filter")
    s_ms, s_n = _hip.profile_get("select")
    w_ms, w_n = _hip.profile_get("walk")
    b_ms, b_n = _hip.profile_get("bounds")
    sc_ms, sc_n = _hip.profile_get("schedule")
    _hip.profile(False)
    f64_check = None
    if args.precision == "f32" or args.filter_precision == "f32":
        # the same segment through the float64 form: the size-independent check of the float32 form
        _hip.call("exm_set_gather_precision", 0)
        _hip.call("exm_set_filter_precision", 0)
        ref64 = pe_intrinsic(seg, bank, st, **kw)
        _hip.call("exm_set_gather_precision", 1 if args.precision == "f32" else 0)
        _hip.call("exm_set_filter_precision", 1 if args.filter_precision == "f32" else 0)
        f64_check = {"max_abs_dln_lambda": float(np.nanmax(np.abs(res["ln_lambda"] - ref64["ln_lambda"]))),
                     "dln_b": float(abs(res["ln_b"][0] - ref64["ln_b"][0]))}

    precision_form = None
    if args.precision == "f64" and args.filter_precision == "f64" and args.config == 2 and not args.no_precision_form:
        # the precision form on the same segment, timed the same way after one warm-up call: float32 transform
        # and interpolation, float64 weighting, index arithmetic, log-sum-exp and prior quadrature
        _hip.call("exm_set_gather_precision", 1)
        _hip.call("exm_set_filter_precision", 1)
        try:
            pe_intrinsic(seg, bank, st, **kw)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                r32 = pe_intrinsic(seg, bank, st, **kw)
            w32 = time.perf_counter() - t0
        finally:
            _hip.call("exm_set_gather_precision", 0)
            _hip.call("exm_set_filter_precision", 0)
        precision_form = {
            "value": args.steps / w32, "unit": "segments/s", "ms_per_step": 1e3 * w32 / args.steps,
            "dtype": "f32 transform and interpolation; f64 weighting, index arithmetic, log-sum-exp and prior quadrature",
            "max_abs_dln_lambda_vs_f64": float(np.nanmax(np.abs(r32["ln_lambda"] - res["ln_lambda"]))),
            "dln_b_vs_f64": float(abs(r32["ln_b"][0] - res["ln_b"][0]))}

    geometry = _node_geometry(bank, st, 0.1267)
    valid, bnode = geometry[0], geometry[3]
    n_nodes_used = int(valid.any(axis=1).sum())
    F = bank.nu.size
    nm, nt = bank.u.size, bank.n_time
    seg_bytes = segment_bytes(F, nm, nt, n_nodes_used)
    g_s = g_ms / 1e3 / max(g_n, 1)                 # one gather launch
    g_step = g_ms / 1e3 / args.steps                # every gather launch of a step (a walk longer than one
    n_jobs = int(valid.sum() + (bnode >= 0).sum())  # launch takes, e.g. every arrival time, runs in slices)
    walk = args.n_loud + args.n_random if args.n_loud > 0 else nt
    lookups = n_jobs * st.tau_idx.size * walk
    achieved = seg_bytes / g_s / 1e9
    value = args.steps / wall
    out = {
        "metric": "segments/sec at ln B err <=1e-5 vs FP64 CPU; p50/p99 latency; HBM GB/s",
        "value": value, "unit": "segments/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64" if args.precision == "f64" and args.filter_precision == "f64"
                  else "f32 %s, f64 elsewhere" % " and ".join(
                      n for n, v in (("transform", args.filter_precision), ("interpolation", args.precision))
                      if v == "f32")),
        "data": "synthetic",
        "config": {"workload": "one 16 s segment, 20-1024 Hz, %s, pe_intrinsic 'volume' prior, %s"
                               % ("2000-template (M_det, q, chi) bank, 53 HBM-resident table nodes" if args.config == 2
                                  else "64-template chirp-mass grid bank, one table node",
                                  "n_loud=%d, n_random=%d" % (args.n_loud, args.n_random) if args.n_loud > 0
                                  else "every arrival time (n_loud=0)"),
                   "n_templates": nm, "n_time": nt, "n_freq": F, "nodes": n_nodes_used, "reads": n_jobs,
                   "injection_snr": round(snr, 3), "device": info["name"], "num_cu": info["num_cu"]},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "gather",
                     "bytes_per_launch": seg_bytes, "kernel_ms": 1e3 * g_s},
        "stages_ms": {"filter": f_ms / args.steps, "select": s_ms / args.steps,
                      "walk": w_ms / args.steps, "bounds": b_ms / args.steps, "schedule": sc_ms / args.steps,
                      "gather": g_ms / args.steps},
        "gather_launches_per_step": g_n / args.steps,
        "lookups_per_step": lookups,
        "lookups_per_s": lookups / g_step,
        "path_hbm_frac": seg_bytes / (wall / args.steps) / 1e9 / HBM_PEAK_GBS,
        "ln_b": float(res["ln_b"][0]),
    }
    trf = measured_traffic(1e3 * g_s)
    if (trf is not None and args.config == 2 and args.grid == "20,10,10" and args.nodes == 53 and walk == 2048
            and args.precision == "f64" and args.gather_kernel in (0, 4) and args.piece in (0, 8)):
        out["roofline"]["traffic"] = trf["traffic_bytes"]
        out["roofline"]["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this kernel on this "
                                             "workload, scaled by a 1 GiB copy (profiles/gather_traffic.json)")
        if "valu_lane_instr_per_lookup" in trf:
            # the gather's actual bound: vector-ALU issue (the measured lane-instructions per lookup at the
            # timed lookup rate, against the issue peak)
            per = trf["valu_lane_instr_per_lookup"]
            out["roofline"]["valu"] = {"lane_instr_per_lookup": per, "achieved": per * lookups / g_s,
                                       "peak": VALU_PEAK_LANE_OPS, "unit": "lane-instructions/s",
                                       "frac": per * lookups / g_s / VALU_PEAK_LANE_OPS,
                                       "source": "SQ_INSTS_VALU x 64 / lookups (profiles/counters_gather_sched_"
                                                 "16s_2000tmpl_53nodes.txt)"}
    if f64_check is not None:
        out["float32_form_vs_float64"] = f64_check
    if precision_form is not None:
        out["precision_form"] = precision_form
    if args.serve_requests > 0 and args.config == 2 and args.nodes == 53:
        # the serving form on the same resident store (config 4's request mix), reported beside the headline
        sv = serve_run(st, args.serve_requests, args.rate, args.seed, args.n_loud, args.n_random)
        sv["workload"] = SERVE_WORKLOAD % args.n_loud
        out["serving"] = sv
    which = [w.strip() for w in args.published.split(",") if w.strip()]
    if which and args.config == 2 and args.nodes == 53 and args.precision == "f64" \
            and args.filter_precision == "f64":
        # the reference's own timed workloads on the same resident store, beside its published times (a
        # failure there is reported in the line, which still carries the headline)
        try:
            out["reference_workloads"] = published_workloads(st, which)
        except Exception as e:                   # noqa: BLE001
            out["reference_workloads"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not args.no_cpu and args.n_loud > 0 and args.config == 2:
        an_geometry = geometry
        cpu, err, nchk = cpu_leg(args, seg, bank, st, res, an_geometry, 0)
        out["cpu_baseline"] = cpu
        out["oracle_check"] = {"templates": nchk, "max_abs_dln_lambda_sample": err,
                               "full_bank": "tests/test_gpu_configs.py::test_config2_full_bank (every template's ln "
                                            "Lambda and ln B against the FP64 oracle)"}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
