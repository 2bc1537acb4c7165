"""Benchmark of the mesh-measurement hot path on one MI355X.

Workload (one step = one mock): 256^3 float32 particles from a Zel'dovich field
(Gaussian field P(k) = 1e4 k^-1.5, L = 1000 Mpc/h, seeded) resident in HBM ->
TSC onto a 512^3 float32 mesh -> r2c FFT -> P(k) in 128 shells (ndel=2) with
MAS deconvolution and shot-noise subtraction -> T-web of the R=2 Mpc/h
Gaussian-smoothed density, lambda_th=0, int8 labels.

Prints one JSON line: mocks/s of the optimised path, p50/p99 latency, the
plain GPU path on the same inputs, the float32-vs-float64 P(k) error of the
measured mock, the roofline of the dominant kernel (HIP events on the
library's stream) and a CPU baseline (the OpenMP restatement of the reference
path in oracle/cpu_ref.c, timed on a bounded sample on the host cores).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TSC+P(k) mocks/s & p50/p99 latency at 512^3 vs plain GPU path, P(k) err≤1e-4"
HBM_PEAK_GBS = 8000.0

NL, N, L = 256, 512, 1000.0
NDEL, R_SMOOTH, LAMBDA_TH = 2, 2.0, 0.0
# T-web margin: count the cells whose label the float32 rounding of the
# potential can change (tol = K(N) u ||phi||_rms / d^2, applied on the device)
MARGIN_TOL = -1.0


MI355X_CUS, MI355X_CLOCK_HZ = 256, 2.4e9


def algorithmic_bytes(name, n_part, n_rec):
    """Bytes a kernel must move once (each array read or written once)."""
    n3 = N ** 3
    nh = N * N * (N // 2 + 1)
    table = {
        "mas_tile_deposit": 12 * n_rec + 4 * n3,
        "mas_tile_scatter": 12 * n_part + 12 * n_rec,
        "mas_tile_count": 12 * n_part,
        "mas_plain": 12 * n_part + 4 * n3,
        "fft_z_r2c": 4 * n3 + 8 * nh,
        "fft_y_fwd": 16 * nh,
        "fft_x_fwd": 16 * nh,
        "fft_x_inv": 16 * nh,
        "fft_y_inv": 16 * nh,
        "fft_z_c2r": 8 * nh + 4 * n3,
        "fused_x": 16 * nh,
        "tweb_fused": 4 * n3 + n3,
        "pk_shells": 8 * nh,
        "tweb_kmul": 16 * nh,
        "tweb_grad": 4 * n3 + 12 * n3,
        "tweb_hess_eig": 12 * n3 + n3,
        "mesh_sum": 4 * n3,
    }
    return table.get(name)


KERNELS = ["mas_group_count", "mas_group_scan", "mas_group_scatter", "mas_tile_count", "mas_tile_prefix", "mas_tile_scan", "mas_tile_items", "mas_tile_scatter", "mas_tile_zero_split",
           "mas_tile_deposit", "mas_plain", "fill", "mesh_sum", "mesh_sum_final", "fft_z_r2c", "fused_sum",
           "fft_y_fwd", "fused_zero", "fused_x", "fft_x_fwd", "pk_zero", "pk_shells", "tweb_kmul", "fft_x_inv",
           "fft_y_inv", "fft_z_c2r", "tweb_grad", "tweb_zero", "tweb_hess_eig", "tweb_fused", "tweb_rare"]


# rocprofv3 kernel names (demangled prefix) -> the library's kernel names, for
# the fused path (the only path run with --fused-only)
ROCPROF_NAMES = [("cm::k_tile_count_lds", "mas_tile_count"), ("cm::k_tile_count", "mas_tile_count"),
                 ("cm::k_tile_prefix", "mas_tile_prefix"), ("cm::k_tile_total", "mas_tile_scan"),
                 ("cm::k_tile_items", "mas_tile_items"), ("cm::k_tile_scatter", "mas_tile_scatter"),
                 ("cm::k_tile_zero_split", "mas_tile_zero_split"), ("cm::k_tile_deposit", "mas_tile_deposit"),
                 ("cm::k_rows_r2c", "fft_z_r2c"), ("cm::k_rows_c2r", "fft_z_c2r"), ("cm::k_x_fused", "fused_x"),
                 ("cm::k_tweb_fused", "tweb_fused"), ("cm::k_tweb_rare", "tweb_rare")]


def lib_name(kname):
    k = kname.replace("void ", "", 1)
    if k.startswith("cm::k_cols<") or k.startswith("cm::k_cols_ct<"):
        targs = [a.strip() for a in k[k.index("<") + 1:k.index(">")].split(",")]
        return "fft_y_fwd" if targs[2] == "true" else "fft_y_inv"
    for pre, name in ROCPROF_NAMES:
        if k.startswith(pre):
            return name
    return None


def _largest_grid(rows):
    """Per library kernel, the rows of its largest grid: the measured mock's
    dispatches, without the smaller ones the input generator makes of the same
    kernels (its 256^3 FFT passes)."""
    big = {}
    for r in rows:
        name = lib_name(r["Kernel_Name"])
        if name:
            big[name] = max(big.get(name, 0), int(r["Grid_Size"]))
    return [r for r in rows if lib_name(r["Kernel_Name"]) and int(r["Grid_Size"]) == big[lib_name(r["Kernel_Name"])]]


def valu_from_pmc(csv_path, out_json):
    """Median SQ_INSTS_VALU per dispatch and library kernel from a rocprofv3
    --pmc pass (SQ_INSTS_VALU SQ_INSTS_LDS SQ_INSTS_SALU SQ_WAVES) of
    `bench.py --fused-only`; the counter is summed over the per-XCD/SE
    instances of each dispatch."""
    import csv
    per = {}
    rows = [r for r in csv.DictReader(open(csv_path)) if r["Counter_Name"] == "SQ_INSTS_VALU"]
    for r in _largest_grid(rows):
        key = (lib_name(r["Kernel_Name"]), r.get("Dispatch_Id", r.get("Correlation_Id", "")))
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    vals = {}
    for (name, _), v in per.items():
        vals.setdefault(name, []).append(v)
    out = {"source": "SQ_INSTS_VALU, wave instructions per dispatch, 512^3 TSC single mock (bench.py --fused-only)",
           "sq_insts_valu": {k: float(np.median(v)) for k, v in sorted(vals.items())}}
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    return out


def traffic_from_pmc(fetch_csv, write_csv, out_json):
    """HBM bytes per dispatch from two rocprofv3 --pmc passes (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch) of `bench.py --fused-only --copy-calibration`.
    Each counter is scaled by what it reads for the 512 MiB device copy made by
    --copy-calibration (the largest __amd_rocclr_copyBuffer dispatch)."""
    import csv
    copy_kib = 4 * N ** 3 / 1024.0

    def load(path, counter):
        vals, copy = {}, []
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        for r in rows:
            if "copyBuffer" in r["Kernel_Name"]:
                copy.append(float(r["Counter_Value"]))
        for r in _largest_grid(rows):
            vals.setdefault(lib_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
        return {k: float(np.median(v)) for k, v in vals.items()}, max(copy)

    fetch, cf = load(fetch_csv, "FETCH_SIZE")
    write, cw = load(write_csv, "WRITE_SIZE")
    ff, fw = copy_kib / cf, copy_kib / cw
    out = {"_factors": {"fetch": round(ff, 4), "write": round(fw, 4), "copy_fetch_kib": cf, "copy_write_kib": cw,
                        "copy_kib": copy_kib}}
    for k in sorted(set(fetch) & set(write)):
        out[k] = int(round((fetch[k] * ff + write[k] * fw) * 1024))
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return out


def cpu_baseline(pos, quick=False):
    """OpenMP restatement of the reference path (oracle/cpu_ref.c) on the host
    cores, measured on the 512^3 mock itself: the same particles, TSC onto
    512^3, r2c, 128 shells, T-web; 1 warm-up and 3 timed runs on all cores of
    the machine's share (16 per GPU), then 3 runs on the reference's 12 OpenMP
    threads (_NTHREADS_).  value = 1 / the median run on all cores.  A cross-check
    scaled from a 256^3 sample (N^3 for the mesh passes, N^3 log N for the
    FFTs, N_p for the deposit) rides along.  quick: the scaled 128^3 sample
    only (iteration runs)."""
    try:
        import oracle.cpu_ref as cr
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "mocks/s", "cores": 0, "kind": "port", "sample": f"unavailable: {e}"}
    threads = int(os.environ.get("CPU_BASELINE_THREADS", min(16, os.cpu_count() or 1)))
    kw = dict(scheme=2, ndel=NDEL, R=R_SMOOTH, lambda_th=LAMBDA_TH)

    def stages(t):
        return {k: round(float(x), 4) for k, x in t.items() if k != "runs"}

    def say(tag):
        return lambda i, sec: print(f"cpu baseline {tag}: run {i + 1} {sec:.2f} s", file=sys.stderr, flush=True)

    def scaled(NS):
        npart = (NS // 2) ** 3
        sub = pos[:: max(1, len(pos) // npart)][:npart].astype(np.float32)
        Ls = L * NS / N  # same cell size d = L/N
        sub = np.mod(sub, Ls).astype(np.float32)
        t = cr.time_pipeline(sub, NS, Ls, threads=threads, repeats=1, warmup=0, progress=say(f"{NS}^3 sample"), **kw)
        f_mesh = (N / NS) ** 3
        f_fft = f_mesh * math.log2(N) / math.log2(NS)
        f_part = len(pos) / npart
        tw_fft = t["tweb_fft"] / max(t["tweb"], 1e-12)
        est = t["mas"] * f_part + t["fft"] * f_fft + t["pk"] * f_mesh + t["tweb"] * (
            f_mesh * tw_fft * math.log2(N) / math.log2(NS) + f_mesh * (1 - tw_fft))
        return {"value": 1.0 / est, "sample": f"{NS}^3 mesh, {npart} particles, 1 run of "
                f"{sum(t[k] for k in ('mas', 'fft', 'pk', 'tweb')):.2f} s, scaled by N^3, N^3 log N and N_p",
                "stage_seconds_sample": stages(t)}

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    host = {"nproc": os.cpu_count(), "model": model}
    fft_note = "the port's own radix-2 CPU FFT (FFTW is not on the host)"
    if quick:
        x = scaled(128)
        return {"value": x["value"], "unit": "mocks/s", "cores": threads, "kind": "port",
                "sample": x["sample"] + "; " + fft_note, "stage_seconds_sample": x["stage_seconds_sample"],
                "host": host}
    full = np.ascontiguousarray(pos, np.float32)
    t = cr.time_pipeline(full, N, L, threads=threads, repeats=3, warmup=1, progress=say(f"{threads} threads"), **kw)
    t12 = cr.time_pipeline(full, N, L, threads=12, repeats=3, warmup=0, progress=say("12 threads"), **kw)
    med = sorted(t["runs"])[1]
    med12 = sorted(t12["runs"])[1]
    cross = scaled(256)
    return {"value": 1.0 / med, "unit": "mocks/s", "cores": threads, "kind": "port",
            "sample": f"the full {N}^3 mock ({len(pos)} particles, TSC, {N // NDEL // 2} shells, T-web): median of "
                      f"3 runs after 1 warm-up, {med:.2f} s per mock; {fft_note}",
            "runs_s": [round(x, 3) for x in t["runs"]], "stage_seconds": stages(t),
            "threads_12": {"value": 1.0 / med12, "runs_s": [round(x, 3) for x in t12["runs"]],
                           "stage_seconds": stages(t12)},
            "scaled_cross_check": cross, "host": host}


def served(cm, posd, b, slots, nreq, rate=None, seed=0):
    """512^3 mocks through the continuous-batching server: `nreq` requests on
    the resident particles, all at once (rate None) or with Poisson arrivals
    at `rate` per second.  Returns mocks/s and latency percentiles (ms)."""
    import threading
    import torch
    rng = np.random.default_rng(seed)
    kw = dict(do_tweb=True, R_smooth=R_SMOOTH, lambda_th=LAMBDA_TH, margin_tol=MARGIN_TOL)
    with cm.Server(0, slots=slots) as srv:
        labels = [torch.empty((N, N, N), dtype=torch.int8, device="cuda") for _ in range(slots + 1)]
        mk = lambda i: srv.request(N, L, "TSC", b["nshells"], b["kmin"], b["dk"], pos=posd,
                                   labels=labels[i % len(labels)], **kw)
        # warm the slots' buffers and graphs: every slot sees every labels
        # buffer (part of a request's graph key) at least twice before timing
        for t in [srv.submit(mk(i)) for i in range(4 * slots * len(labels))]:
            t.result()
        torch.cuda.synchronize()
        tickets = [None] * nreq
        t0 = time.perf_counter()
        if rate is None:
            for i in range(nreq):
                tickets[i] = srv.submit(mk(i))
        else:
            def producer():
                nxt = time.perf_counter()
                for i in range(nreq):
                    nxt += float(rng.exponential(1.0 / rate))
                    d = nxt - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    tickets[i] = srv.submit(mk(i))
            th = threading.Thread(target=producer)
            th.start()
            th.join()
        lat = np.array([t.result()["latency_ms"] for t in tickets])
        T = time.perf_counter() - t0
        adm = srv.stats()
    return {"value": round(nreq / T, 2), "unit": "mocks/s", "slots": slots, "requests": nreq,
            "arrivals": "all at t=0" if rate is None else f"Poisson, {rate:.0f}/s",
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "admission": adm}


MIXED_NL = (48, 64, 96, 128, 192, 256, 384)  # FFT sizes for the generator's Lagrangian grid


def served_mixed(cm, slots=3, nreq=96, rate=None, seed=5, specs=None):
    """Mixed serving workload: N uniform in {128, 192, 256, 384, 512, 768}, NL^3
    particles with NL^3 log-uniform in [1e5, 1e8] and NL rounded (in log) to the
    nearest Lagrangian grid the generator's FFT takes (48 .. 384), scheme uniform in NGP/CIC/TSC,
    T-web on half of the requests, each mock generated on the GPU from its own
    seed; admission by HBM footprint.  All requests at once (rate None) or Poisson
    arrivals at `rate` per second.  Returns mocks/s, latency percentiles (ms) and
    the specs (so the Poisson leg replays the same requests)."""
    import threading
    import torch
    rng = np.random.default_rng(seed)
    if specs is None:
        specs = []
        for i in range(nreq):
            Nm = int(rng.choice([128, 192, 256, 384, 512, 768]))
            nl = math.exp(rng.uniform(math.log(1e5), math.log(1e8))) ** (1.0 / 3.0)
            NLm = min(MIXED_NL, key=lambda g: abs(math.log(g / nl)))
            Lm = float(rng.uniform(250.0, 2000.0))
            sch = ["NGP", "CIC", "TSC"][int(rng.integers(0, 3))]
            tw = bool(rng.integers(0, 2))
            specs.append((Nm, NLm, Lm, sch, tw, 2000 + i))
    nreq = len(specs)

    def mk(spec):
        Nm, NLm, Lm, sch, tw, sd = spec
        ndel = 1 if Nm in (128, 192) else Nm // 128
        bm = cm.binning(Nm, Lm, ndel=ndel)
        kw = dict(do_tweb=True, R_smooth=2.0 * Lm / Nm, margin_tol=MARGIN_TOL) if tw else {}
        return srv.request(Nm, Lm, sch, 64, bm["kmin"], bm["dk"], gen=dict(NL=NLm, seed=sd), **kw)

    with cm.Server(0, slots=slots) as srv:
        for t in [srv.submit(mk(sp)) for sp in specs]:  # warm-up: buffers, tables, graphs
            t.result()
        torch.cuda.synchronize()
        tickets = [None] * nreq
        t0 = time.perf_counter()
        if rate is None:
            for i in range(nreq):
                tickets[i] = srv.submit(mk(specs[i]))
        else:
            def producer():
                nxt = time.perf_counter()
                for i in range(nreq):
                    nxt += float(rng.exponential(1.0 / rate))
                    d = nxt - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    tickets[i] = srv.submit(mk(specs[i]))
            th = threading.Thread(target=producer)
            th.start()
            th.join()
        lat = np.array([t.result()["latency_ms"] for t in tickets])
        T = time.perf_counter() - t0
        adm = srv.stats()
    npart = sum(sp[1] ** 3 for sp in specs)
    return {"value": round(nreq / T, 2), "unit": "mocks/s", "slots": slots, "requests": nreq,
            "particles": int(npart),
            "workload": "N in {128..768}, NL^3 particles log-uniform 1e5-1e8 (NL rounded to 48..384) "
                        "generated on the GPU, NGP/CIC/TSC, 64 shells, T-web on half",
            "arrivals": "all at t=0" if rate is None else f"Poisson, {rate:.0f}/s",
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "admission": adm}, specs


def batched(cm, nmocks=64, slots=3, seed=7):
    """64 uneven mocks (N in {128,192,256,384,512}, NL^3 particles with NL in
    {48..384} ~ log-uniform 1e5..5e7, L in [250,2000], NGP/CIC/TSC, 64 shells,
    T-web labels), each generated on the GPU from its own seed, through
    cm_measure_batch.  N=192 uses ndel=1 and keeps 64 of its 96 shells."""
    import torch
    rng = np.random.default_rng(seed)
    Ns = [128, 192, 256, 384, 512]
    NLs = [48, 64, 96, 128, 192, 256, 384]
    specs = []
    for i in range(nmocks):
        Nm = int(rng.choice(Ns))
        NLm = int(rng.choice(NLs))
        Lm = float(rng.uniform(250.0, 2000.0))
        sch = ["NGP", "CIC", "TSC"][int(rng.integers(0, 3))]
        ndel = max(1, Nm // 128) if Nm != 192 else 1
        specs.append((Nm, NLm, Lm, sch, ndel, 1000 + i))
    npart = sum(s[1] ** 3 for s in specs)
    # launch shapes per mesh size from the autotuner (process-wide table the
    # server's slot contexts consult), tuned on a TSC + T-web mock of that N
    tuned = {}
    with cm.Context(0) as tctx:
        for Nm in Ns:
            NLm = {128: 64, 192: 96, 256: 128, 384: 192, 512: 256}[Nm]
            Lm = 1000.0
            posm = tctx.zeldovich(NLm, Lm, A=1e4, ns=-1.5, seed=77)
            bm = cm.binning(Nm, Lm, ndel=max(1, Nm // 128) if Nm != 192 else 1)
            r = tctx.autotune(posm, Nm, Lm, "TSC", 64, bm["kmin"], bm["dk"], do_tweb=True, R_smooth=2.0 * Lm / Nm,
                              margin_tol=MARGIN_TOL)
            tuned[str(Nm)] = {"tune": r["tune"], "kernel_ms_before": round(r["ms_before"], 3),
                              "kernel_ms_after": round(r["ms_after"], 3)}
            del posm
        torch.cuda.synchronize()
    with cm.Server(0, slots=slots) as srv:
        def tickets():
            out = []
            for (Nm, NLm, Lm, sch, ndel, sd) in specs:
                bm = cm.binning(Nm, Lm, ndel=ndel)
                out.append(srv.request(Nm, Lm, sch, 64, bm["kmin"], bm["dk"], gen=dict(NL=NLm, seed=sd),
                                       do_tweb=True, R_smooth=2.0 * Lm / Nm, margin_tol=MARGIN_TOL))
            return out
        srv.batch(tickets()[:slots * 2])  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = srv.batch(tickets())
        T = time.perf_counter() - t0
        adm = srv.stats()
    lat = np.array([r["latency_ms"] for r in res])
    # the same 64 mocks with their particles generated beforehand: through the
    # slot server, and through cm_measure_packed (one deposit launch per scheme
    # over the union (mock, tile) work list, one T-web launch for all mocks)
    with cm.Context(0) as pctx:
        args = []
        for (Nm, NLm, Lm, sch, ndel, sd) in specs:
            bm = cm.binning(Nm, Lm, ndel=ndel)
            args.append(dict(pos=pctx.zeldovich(NLm, Lm, A=1e4, ns=-1.5, seed=sd), N=Nm, L=Lm, scheme=sch,
                             nshells=64, kmin=bm["kmin"], dk_bin=bm["dk"], do_tweb=True, R_smooth=2.0 * Lm / Nm,
                             margin_tol=MARGIN_TOL))
        torch.cuda.synchronize()
        with cm.Server(0, slots=slots) as srv:
            def pre():
                return [srv.request(a["N"], a["L"], a["scheme"], 64, a["kmin"], a["dk_bin"], pos=a["pos"],
                                    do_tweb=True, R_smooth=a["R_smooth"], margin_tol=MARGIN_TOL) for a in args]
            srv.batch(pre()[:slots * 2])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            srv.batch(pre())
            T_pre = time.perf_counter() - t0
        pctx.measure_packed(args)  # warm-up: buffers and tables
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pctx.measure_packed(args)
        torch.cuda.synchronize()
        T_packed = time.perf_counter() - t0
        del args
    torch.cuda.empty_cache()
    pregen = {"server": {"value": round(nmocks / T_pre, 2), "seconds": round(T_pre, 3), "slots": slots},
              "packed": {"value": round(nmocks / T_packed, 2), "seconds": round(T_packed, 3),
                         "entry": "cm_measure_packed"},
              "note": "particles generated on the GPU before the timed region; the packed path runs the deposits "
                      "of all mocks as one launch per MAS scheme over the union (mock, tile) work list and the "
                      "T-web of all mocks as one launch"}
    return {"value": round(nmocks / T, 2), "unit": "mocks/s", "mocks": nmocks, "slots": slots,
            "particles": int(npart), "seconds": round(T, 3),
            "workload": "64 uneven mocks generated on the GPU (N 128-512, NL^3 particles 1.1e5-5.7e7, "
                        "NGP/CIC/TSC, 64 shells, T-web)",
            "autotune": tuned, "graphs": "hipGraph replay of repeated request shapes per slot",
            "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "admission": adm, "pregenerated": pregen}


ROW_KERNELS = KERNELS + ["zel_spectra", "zel_displace", "uniform", "fkp_wscale", "fkp_wsums", "bk_masks", "bk_sums",
                         "bk_reduce", "bkd_table", "bkd_list", "bkd_pairs", "fft_x_inv", "fft_x_fwd"]
FP64_PEAK_TFS = 78.6


def _timed(fn, reps, warm=2):
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def _kernels(ctx, fn, nrep=3):
    """Mean HIP-event time and launches per call of every library kernel fn runs."""
    import torch
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(nrep):
        fn()
    torch.cuda.synchronize()
    out = {}
    for k in ROW_KERNELS:
        ms, cnt = ctx.kernel_time(k)
        if cnt:
            out[k] = {"mean_ms": ms, "launches_per_call": cnt / nrep}
    ctx.profile(False)
    return out


def _hbm_roof(kname, ktimes, nbytes):
    ms = ktimes[kname]["mean_ms"]
    ach = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": kname, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "algorithmic_bytes": int(nbytes),
            "mean_ms": round(ms, 4)}


def rows(cm):
    """The rows either side of the measurement path, each timed on the GPU with
    the roofline of its dominant kernel and a CPU leg on a bounded sample:
    F1 the Zel'dovich generator (256^3 particles), F2 the FKP estimator with a
    random catalogue (256^3 Zel'dovich data + 4x uniform randoms, TSC 512^3),
    F4 the shell-transform bispectrum (512^3 mesh, 16 shells of 4 k_f) and the
    direct bispectrum (64^3 mesh, 32 bins of k_f, all ordered pairs)."""
    import torch
    import oracle
    import oracle.zeldovich as oz
    res = {}
    ctx = cm.Context(0)
    # ---------------------------------------------------------------- F1
    NLg, Lg = 256, 1000.0
    buf = torch.empty((NLg ** 3, 3), dtype=torch.float32, device="cuda")
    gen = lambda: ctx.zeldovich(NLg, Lg, A=1e4, ns=-1.5, seed=1, out=buf)
    t = _timed(gen, 10)
    kt = _kernels(ctx, gen)
    nh = NLg * NLg * (NLg // 2 + 1)
    ncpu = 64
    tc = time.perf_counter()
    oz.zeldovich(ncpu, Lg, seed=1)
    tcpu = (time.perf_counter() - tc) * (NLg / ncpu) ** 3 * math.log2(NLg) / math.log2(ncpu)
    res["F1_generator"] = {
        "workload": "256^3 Zel'dovich particles (fixed-amplitude field P = 1e4 k^-1.5, L = 1000, Philox phases, three c2r)",
        "value": round(NLg ** 3 / t, 1), "unit": "particles/s", "ms_per_call": round(1e3 * t, 4),
        "kernels_ms_per_call": {k: round(v["mean_ms"] * v["launches_per_call"], 4) for k, v in kt.items()},
        "roofline": _hbm_roof("zel_spectra", kt, 24 * nh),
        "roofline_note": "zel_spectra writes three complex64 spectra (24 B per half-spectrum mode); it is bound by "
                         "the Philox draws and FP64 trigonometry, not by HBM",
        "cpu_baseline": {"value": round(NLg ** 3 / tcpu, 1), "unit": "particles/s", "cores": 1, "kind": "port",
                         "sample": f"numpy restatement (oracle/zeldovich.py) at {ncpu}^3, "
                                   f"{tcpu * (ncpu / NLg) ** 3 * math.log2(ncpu) / math.log2(NLg):.2f} s, "
                                   "scaled by N^3 log N"}}
    del buf
    # ---------------------------------------------------------------- F2
    posd = ctx.zeldovich(NL, L, A=1e4, ns=-1.5, seed=1)
    nr = 4 * NL ** 3
    rand = ctx.uniform(nr, L, seed=2)
    b = cm.binning(N, L, ndel=NDEL, DeltaKmin=0.0)
    nbar = NL ** 3 / L ** 3
    fkp = lambda: ctx.measure_fkp(posd, rand, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], nbar=nbar,
                                  P_est=1e4, fkp_weight=True)
    t = _timed(fkp, 5)
    kt = _kernels(ctx, fkp)
    dom = max(kt.items(), key=lambda kv: kv[1]["mean_ms"] * kv[1]["launches_per_call"])[0]
    # a MAS kernel runs once on the data and once per x-slab group of the
    # randoms: algorithmic bytes of the mean particle count per launch
    n_launch = (NL ** 3 + nr) / max(1.0, kt[dom]["launches_per_call"])
    ab = algorithmic_bytes(dom, n_launch, 1.42 * n_launch)
    # CPU: the oracle's serial MAS, FFT and shells on a 128^3 sample with 1/64 of both catalogues
    Ns = 128
    sub = posd[:: 64].cpu().numpy()
    rsub = rand[:: 64].cpu().numpy()
    Lsub = L * Ns / N
    sub = np.mod(sub * np.float32(Ns / N), Lsub).astype(np.float32)
    rsub = np.mod(rsub * np.float32(Ns / N), Lsub).astype(np.float32)
    tc = time.perf_counter()
    m1 = oracle.mas(sub, Ns, Lsub, oracle.TSC)
    m2 = oracle.mas(rsub, Ns, Lsub, oracle.TSC)
    dk = oracle.fft_r2c(m1 - m2 * (len(sub) / len(rsub)))
    o = oracle.pk_opts(Ns, Lsub, Ns // NDEL // 2, 0.0, NDEL * 2 * np.pi / Lsub, oracle.TSC, 1.0, 1.0)
    oracle.pk_shells(dk, o)
    tcs = time.perf_counter() - tc
    tcpu = tcs * max(64.0, (N / Ns) ** 3 * math.log2(N) / math.log2(Ns))
    res["F2_fkp_randoms"] = {
        "workload": f"256^3 Zel'dovich data + {nr} uniform randoms, FKP weights (P_est = 1e4), TSC 512^3, "
                    "128 shells, monopole/quadrupole/hexadecapole",
        "value": round(1.0 / t, 3), "unit": "measurements/s", "ms_per_call": round(1e3 * t, 4),
        "kernels_ms_per_call": {k: round(v["mean_ms"] * v["launches_per_call"], 4) for k, v in kt.items()},
        "roofline": _hbm_roof(dom, kt, ab) if ab else None,
        "roofline_note": "per launch of the dominant kernel (the data and each x-slab group of the randoms launch "
                         "it once) at the mean particle count per launch, 1.42 records per particle for TSC on 16^3 "
                         "tiles",
        "cpu_baseline": {"value": round(1.0 / tcpu, 5), "unit": "measurements/s", "cores": 1, "kind": "port",
                         "sample": f"the oracle's serial MAS of both catalogues, r2c and shells at {Ns}^3 with 1/64 "
                                   f"of the particles, {tcs:.2f} s, scaled by the larger of the particle (x64) and "
                                   "the N^3 log N mesh factor"}}
    # ---------------------------------------------------------------- F4 fast
    out = ctx.measure(posd, N, L, "TSC", b["nshells"], b["kmin"], b["dk"])
    mesh = ctx.mesh_assign(posd, N, L, "TSC")
    nobj = float(out["n_obj"])
    dkd = ctx.fft_r2c(mesh, subtract=nobj / N ** 3)
    kf = 2 * np.pi / L
    nsb = 16
    bkf = lambda: ctx.bispectrum(dkd, N, L, nsb, 4 * kf, 4 * kf, "TSC", nobj)
    t = _timed(bkf, 3, warm=1)
    kt = _kernels(ctx, bkf, nrep=2)
    r = bkf()
    res["F4_bispectrum_fast"] = {
        "workload": f"512^3 TSC mesh of the 256^3 mock, {nsb} shells of 4 k_f from 4 k_f, "
                    f"{len(r['bk'])} triangle bins, shell transforms",
        "value": round(1.0 / t, 3), "unit": "bispectra/s", "ms_per_call": round(1e3 * t, 3),
        "kernels_ms_per_call": {k: round(v["mean_ms"] * v["launches_per_call"], 4) for k, v in kt.items()},
        "roofline": _hbm_roof("bk_sums", kt, 3 * nsb * N ** 3 * 4),
        "roofline_note": "bk_sums reads the three float32 fields of every shell once (12 ns N^3 bytes); it is bound "
                         "by VALU issue, not HBM: bins in tiles of 2 j x 4 k shells, 8 VALU per cell and tile slot",
        "cpu_baseline": None,
        "cpu_note": "no CPU restatement of the shell-transform estimator; the direct estimator below carries the CPU leg"}
    del mesh, dkd
    # ---------------------------------------------------------------- F4 direct
    Nd, nbd = 64, 32
    Ld = 1000.0
    kfd = 2 * np.pi / Ld
    pd = ctx.zeldovich(64, Ld, A=1e4, ns=-1.5, seed=3)
    md = ctx.mesh_assign(pd, Nd, Ld, "TSC")
    nd = float(md.double().sum())
    dkk = ctx.fft_r2c(md, subtract=nd / Nd ** 3)
    bkd = lambda: ctx.bispectrum_direct(dkk, Nd, Ld, nbd, kfd, kfd, "TSC", nd)
    t = _timed(bkd, 3, warm=1)
    kt = _kernels(ctx, bkd, nrep=2)
    rd = bkd()
    pairs = int(rd["ntri"].sum())
    flops = 11.0 * pairs  # Re(d1 d2 d3): two complex products, real part
    ach = flops / (kt["bkd_pairs"]["mean_ms"] * 1e-3) / 1e12
    Nc = 12
    dkc = oracle.fft_r2c(np.random.default_rng(0).standard_normal((Nc, Nc, Nc)))
    tc = time.perf_counter()
    oracle.bk_direct(dkc, 1, 1, Nc)
    tcs = time.perf_counter() - tc
    cpu_pairs_s = float((Nc + 1) ** 6) / tcs
    tcpu = float((Nd + 1) ** 6) / cpu_pairs_s
    res["F4_bispectrum_direct"] = {
        "workload": f"64^3 TSC mesh of a 64^3 mock, {nbd} bins of k_f from k_f, {pairs} ordered pairs",
        "value": round(1.0 / t, 3), "unit": "bispectra/s", "ms_per_call": round(1e3 * t, 3),
        "pairs_per_s": round(pairs / t, 1),
        "kernels_ms_per_call": {k: round(v["mean_ms"] * v["launches_per_call"], 4) for k, v in kt.items()},
        "roofline": {"kernel": "bkd_pairs", "bound": "fp64", "achieved": round(ach, 3), "peak": FP64_PEAK_TFS,
                     "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFS, 4), "traffic": None,
                     "mean_ms": round(kt["bkd_pairs"]["mean_ms"], 4)},
        "roofline_note": "11 FP64 flops per pair; the kernel is bound by the LDS adds of its run-length sums, the "
                         "table reads hit L1/L2",
        "cpu_baseline": {"value": round(1.0 / tcpu, 8), "unit": "bispectra/s", "cores": 1, "kind": "port",
                         "sample": f"the oracle's loop over all (N+1)^6 ordered pairs (the reference's) at N = {Nc}: "
                                   f"{tcs:.2f} s, {cpu_pairs_s:.3e} pairs/s, scaled to the (N+1)^6 pairs of N = {Nd}"}}
    ctx.close()
    return {"metric": "rows either side of the measurement path on one MI355X", "n_gpus": 1, "rows": res}


def large(cm, steps=3):
    """Large single mock: 5.4e8 Zel'dovich particles (every other Lagrangian x
    plane of a 1024^3 field, P = 1e4 k^-1.5, L = 2000) plus 20% uniform
    randoms (1.07e8), all float32 and generated on the GPU, interlaced TSC onto
    1024^3, 256 shells (ndel = 2).  Reports the time per measurement and the
    per-shell float32-vs-FP64 error of the raw power (the FP64 leg: two FP64
    meshes, FP64 FFTs, combination and FP64 shells)."""
    import torch
    NLg, Nl, Ll = 1024, 1024, 2000.0
    ctx = cm.Context(0)
    zel = ctx.zeldovich(NLg, Ll, A=1e4, ns=-1.5, seed=11)
    n1 = NLg ** 3 // 2
    n2 = n1 // 5
    pos = torch.empty((n1 + n2, 3), dtype=torch.float32, device="cuda")
    pos[:n1] = zel.view(NLg, NLg * NLg, 3)[::2].reshape(-1, 3)
    del zel
    torch.cuda.empty_cache()
    ctx.uniform(n2, Ll, seed=12, out=pos[n1:])
    torch.cuda.synchronize()
    bl = cm.binning(Nl, Ll, ndel=2)
    kw = dict(interlace=True)
    out = ctx.measure(pos, Nl, Ll, "TSC", bl["nshells"], bl["kmin"], bl["dk"], **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        out = ctx.measure(pos, Nl, Ll, "TSC", bl["nshells"], bl["kmin"], bl["dk"], **kw)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    kt = _kernels(ctx, lambda: ctx.measure(pos, Nl, Ll, "TSC", bl["nshells"], bl["kmin"], bl["dk"], **kw), nrep=1)
    raw32 = ctx.measure(pos, Nl, Ll, "TSC", bl["nshells"], bl["kmin"], bl["dk"], sn_correction=False, **kw)
    # FP64: both meshes, FFTs and the interlaced combination in double
    nobj = float(n1 + n2)
    V = Ll ** 3
    specs = []
    for sh in (0.0, 0.5):
        m = ctx.mesh_assign(pos, Nl, Ll, "TSC", interlace_shift=sh, mesh_dtype=torch.float64)
        specs.append(ctx.fft_r2c(m, subtract=nobj / Nl ** 3))
        del m
    j = torch.arange(Nl, device="cuda", dtype=torch.float64)
    sj = torch.where(j <= Nl // 2, j, j - Nl)
    kz = torch.arange(Nl // 2 + 1, device="cuda", dtype=torch.float64)
    a = specs[1]
    for ax, v in ((0, sj), (1, sj), (2, kz)):
        shape = [1, 1, 1]
        shape[ax] = -1
        a.mul_(torch.polar(torch.ones_like(v), torch.pi * v / Nl).view(shape))
    specs[0].add_(a).mul_(0.5)
    del a, specs[1]
    raw64, _, _, nm64 = ctx.pk_shells(specs[0], Nl, Ll, bl["nshells"], bl["kmin"], bl["dk"], "TSC", 0.0, nobj ** 2 / V)
    del specs
    torch.cuda.empty_cache()
    ok = nm64 > 0
    err = np.abs(raw32["pk0"][ok] - raw64[ok]) / np.abs(raw64[ok])
    ctx.close()
    return {"metric": "interlaced TSC 1024^3 P(k) measurement", "value": round(1e3 / ms, 3), "unit": "mocks/s",
            "ms_per_mock": round(ms, 2), "particles": n1 + n2,
            "config": {"workload": "5.4e8 Zel'dovich particles (every other x plane of 1024^3, P=1e4 k^-1.5, "
                                   "L=2000) + 20% uniform randoms, float32, generated on the GPU; "
                                   "interlaced TSC 1024^3; 256 shells"},
            "pk_rel_err_fp32_vs_fp64_max": float(err.max()), "pk_rel_err_fp32_vs_fp64_median": float(np.median(err)),
            "pk_rel_err_per_shell": [float(f"{e:.3g}") for e in err],
            "nmodes_equal_fp64": bool(np.array_equal(nm64, raw32["nmodes"])),
            "kernels_ms_per_mock": {k: round(v["mean_ms"] * v["launches_per_call"], 3) for k, v in kt.items()}}


def run_steps(ctx, posd, b, fused, steps, warmup, labels):
    import torch
    kw = dict(do_tweb=True, R_smooth=R_SMOOTH, lambda_th=LAMBDA_TH, margin_tol=MARGIN_TOL, fused=fused, labels=labels)
    for _ in range(warmup):
        out = ctx.measure(posd, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], **kw)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        out = ctx.measure(posd, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], **kw)
        lat.append(time.perf_counter() - s)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return out, total, np.array(lat)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--quick-cpu", action="store_true", help="smaller CPU sample (128^3)")
    ap.add_argument("--plain-steps", type=int, default=None)
    ap.add_argument("--fused-only", action="store_true",
                    help="time only the optimised path (no plain path, FP64 check or CPU leg): for profiler runs")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy a 512 MiB device buffer once before the steps (reference for PMC byte counters)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served and batched legs")
    ap.add_argument("--rows", action="store_true",
                    help="time the generator, FKP-with-randoms and bispectrum rows instead (one JSON line)")
    ap.add_argument("--large", action="store_true",
                    help="run the large interlaced 1024^3 configuration instead (one JSON line)")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="convert two rocprofv3 --pmc counter files into profiles/traffic.json and exit")
    ap.add_argument("--valu-from", metavar="VALU_CSV",
                    help="convert a rocprofv3 --pmc SQ_INSTS_VALU counter file into profiles/valu.json and exit")
    args = ap.parse_args()
    if args.valu_from:
        print(json.dumps(valu_from_pmc(args.valu_from, os.path.join(ROOT, "profiles", "valu.json"))))
        return
    if args.traffic_from:
        print(json.dumps(traffic_from_pmc(args.traffic_from[0], args.traffic_from[1],
                                          os.path.join(ROOT, "profiles", "traffic.json"))))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import cosmicatlas_amd as cm

    if args.large:
        print(json.dumps(large(cm)))
        return
    if args.rows:
        print(json.dumps(rows(cm)))
        return
    dev = torch.device("cuda:0")
    ctx = cm.Context(0)
    # synthetic input generated on the GPU (seeded Gaussian field + Zel'dovich)
    torch.cuda.synchronize()
    t = time.time()
    posd = ctx.zeldovich(NL, L, A=1e4, ns=-1.5, seed=1)
    torch.cuda.synchronize()
    gen_s = time.time() - t
    pos = posd.cpu().numpy()
    b = cm.binning(N, L, ndel=NDEL, DeltaKmin=0.0)
    labels = torch.empty((N, N, N), dtype=torch.int8, device=dev)
    if args.copy_calibration:
        xa = torch.ones(N ** 3, dtype=torch.float32, device=dev)
        xb = torch.empty_like(xa)
        xb.copy_(xa)
        torch.cuda.synchronize()
        del xa, xb

    # launch shapes from the autotuner (kernel time per mock before/after); the
    # profiler runs (--fused-only) keep the built-in shapes so that every
    # dispatch of a kernel in the trace has the same shape
    tune = None
    if not args.fused_only:
        tune = ctx.autotune(posd, N, L, "TSC", b["nshells"], b["kmin"], b["dk"], do_tweb=True, R_smooth=R_SMOOTH,
                            lambda_th=LAMBDA_TH, margin_tol=MARGIN_TOL, labels=labels)
    # optimised path (hipGraph replay after the first two steps)
    out, total, lat = run_steps(ctx, posd, b, True, args.steps, args.warmup, labels)
    labels_fused = labels.clone()  # the plain path below reuses the buffer
    graph_counts = ctx.graph_counts()
    value = args.steps / total
    # per-kernel times (HIP events on the library stream) over a few more steps
    ctx.profile(True)
    ctx.profile_reset()
    nprof = 3
    run_steps(ctx, posd, b, True, nprof, 0, labels)
    ktimes = {}
    for k in KERNELS:
        ms, cnt = ctx.kernel_time(k)
        if cnt:
            ktimes[k] = {"mean_ms": ms, "launches_per_step": cnt / nprof}
    ctx.profile(False)
    n_rec = None
    # dominant kernel: largest total time per step
    dom = max(ktimes.items(), key=lambda kv: kv[1]["mean_ms"] * kv[1]["launches_per_step"])
    # records per deposit: estimated from the ghost fraction measured on the host
    d = L / N
    c = np.floor(pos.astype(np.float64) / d).astype(np.int64) % 16
    n_rec = int(np.prod(1 + (c == 0) + (c == 15), axis=1).sum())
    ab = algorithmic_bytes(dom[0], len(pos), n_rec)
    roof = None
    if ab is not None:
        ach = ab / (dom[1]["mean_ms"] * 1e-3) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tf):
            traffic = json.load(open(tf)).get(dom[0])
        roof = {"kernel": dom[0], "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "algorithmic_bytes": ab,
                "mean_ms": round(dom[1]["mean_ms"], 4)}
        vf = os.path.join(ROOT, "profiles", "valu.json")
        insts = json.load(open(vf))["sq_insts_valu"].get(dom[0]) if os.path.exists(vf) else None
        if insts and N == 512:
            # a wave64 VALU instruction holds one of the 4 SIMDs of a CU for 4 cycles
            bound_ms = insts * 4 / (MI355X_CUS * 4) / MI355X_CLOCK_HZ * 1e3
            roof["valu_issue"] = {"sq_insts_valu": insts, "bound_ms": round(bound_ms, 4),
                                  "frac": round(bound_ms / dom[1]["mean_ms"], 3),
                                  "note": "instruction count from a rocprofv3 SQ_INSTS_VALU pass of this workload; "
                                          "frac = issue-bound time / measured mean"}
            # the roof that binds: the larger of the VALU issue time and the
            # time of the kernel's measured HBM traffic at peak bandwidth
            hbm_ms = (traffic or ab) / (HBM_PEAK_GBS * 1e9) * 1e3
            if bound_ms > hbm_ms:
                roof["bound"] = "valu"
                roof["note"] = ("bound by VALU issue (valu_issue.bound_ms > the HBM time of its traffic, "
                                f"{hbm_ms:.4f} ms); achieved/peak/frac are its algorithmic bytes against HBM peak")

    if args.fused_only:
        print(json.dumps({"metric": METRIC, "value": round(value, 3), "unit": "mocks/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(1e3 * total / args.steps, 3), "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "fused path only (profiler run)", "N": N},
                          "kernels_ms_per_step": {k: round(v["mean_ms"] * v["launches_per_step"], 4)
                                                  for k, v in ktimes.items()}, "roofline": roof}))
        ctx.close()
        return

    # plain GPU path on the same inputs
    psteps = args.plain_steps or max(3, args.steps // 2)
    pout, ptotal, plat = run_steps(ctx, posd, b, False, psteps, 1, labels)
    plain_value = psteps / ptotal

    # float32 vs float64 P(k) of this mock: float64 mesh, FFT and shells on the GPU.
    # Errors are taken relative to the un-subtracted power |P| + SN-term (the
    # quantity the FFT computes; at high k this field's P(k) is ~0 after the
    # shot noise is removed), and separately over shells where P > SN.
    mesh64 = ctx.mesh_assign(posd, N, L, "TSC", mesh_dtype=torch.float64)
    nobj = ctx.mesh_sum(mesh64)
    dk64 = ctx.fft_r2c(mesh64, subtract=nobj / N ** 3)
    del mesh64
    V = L ** 3
    r0, _, _, rn = ctx.pk_shells(dk64, N, L, b["nshells"], b["kmin"], b["dk"], "TSC", V / nobj, nobj ** 2 / V)
    raw0 = ctx.pk_shells(dk64, N, L, b["nshells"], b["kmin"], b["dk"], "TSC", 0.0, nobj ** 2 / V)[0]
    del dk64
    torch.cuda.empty_cache()
    valid = rn > 0
    err = np.abs(out["pk0"][valid] - r0[valid])
    pk_err = float(np.max(err / np.abs(raw0[valid])))
    sig = valid & (r0 > V / nobj)
    pk_err_sig = float(np.max(np.abs(out["pk0"][sig] - r0[sig]) / r0[sig])) if sig.any() else None
    modes_equal = bool(np.array_equal(rn, out["nmodes"]))
    # T-web labels of the timed path against an FP64 T-web of the same particles
    # (FP64 deposit, FFTs, FD Hessian, eigenvalues): a label may differ only
    # where an FP64 eigenvalue lies within the applied tolerance of lambda_th
    m64 = ctx.mesh_assign(posd, N, L, "TSC", algo="plain", mesh_dtype=torch.float64)
    ctx.mesh_affine(m64, N ** 3 / len(pos), -1.0)
    lab64, lam64, _, _ = ctx.tweb_f64(m64, L, R_SMOOTH, LAMBDA_TH, margin_tol=0.0, want_lambdas=True)
    del m64
    marg64 = (lam64 - LAMBDA_TH).abs().min(dim=0).values
    del lam64
    differ = labels_fused != lab64
    tweb_check = {"labels_differing_from_fp64": int(differ.sum()),
                  "all_differing_inside_band": bool((marg64[differ] < out["tweb_tol"]).all()),
                  "margin_tol": out["tweb_tol"], "counted_marginal": out["tweb_marginal"]}
    del lab64, marg64, differ
    torch.cuda.empty_cache()

    serve = None
    if not args.no_serve:
        closed = served(cm, posd, b, slots=2, nreq=40)
        poisson = served(cm, posd, b, slots=2, nreq=60, rate=0.8 * closed["value"], seed=1)
        mixed, specs = served_mixed(cm)
        mixed_p, _ = served_mixed(cm, rate=0.8 * mixed["value"], seed=6, specs=specs)
        serve = {"closed_loop": closed, "poisson": poisson, "batched_c4": batched(cm),
                 "mixed_c5": {"closed_loop": mixed, "poisson": mixed_p}}

    cpu = None if args.no_cpu else cpu_baseline(pos, quick=args.quick_cpu)

    res = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "mocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single mock: 256^3 Zel'dovich float32 particles (P=1e4 k^-1.5, L=1000 Mpc/h, "
                               "seed 1, generated on the GPU) -> "
                               "TSC 512^3 float32 mesh -> r2c -> P(k) 128 shells (ndel=2, FP64 sums) -> "
                               "T-web (Gaussian R=2 Mpc/h, lambda_th=0, int8 labels)",
                   "N": N, "particles": int(len(pos)), "L": L, "nshells": b["nshells"]},
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "plain_gpu": {"value": round(plain_value, 3), "unit": "mocks/s",
                      "p50_ms": round(1e3 * float(np.percentile(plat, 50)), 3),
                      "p99_ms": round(1e3 * float(np.percentile(plat, 99)), 3)},
        "speedup_vs_plain": round(value / plain_value, 3),
        "pk_rel_err_vs_fp64": pk_err,
        "pk_rel_err_vs_fp64_note": "max over shells of |P32-P64| / (P64 before shot-noise subtraction)",
        "pk_rel_err_signal_shells": pk_err_sig,
        "signal_shells": int(sig.sum()),
        "nmodes_equal_fp64": modes_equal,
        "tweb_counts": out["tweb_counts"],
        "tweb_marginal": out["tweb_marginal"],
        "tweb_vs_fp64": tweb_check,
        "kernels_ms_per_step": {k: round(v["mean_ms"] * v["launches_per_step"], 4) for k, v in ktimes.items()},
        "roofline": roof,
        "autotune": {"tune": tune["tune"], "kernel_ms_before": round(tune["ms_before"], 4),
                     "kernel_ms_after": round(tune["ms_after"], 4)},
        "graphs": {"captured": graph_counts[0], "launched": graph_counts[1]},
        "served": serve,
        "cpu_baseline": cpu,
        "input_gen_ms": round(1e3 * gen_s, 2),
    }
    if cpu and cpu.get("value"):
        res["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    ctx.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
