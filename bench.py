"""Benchmark of the heat-bath VCI engine on one MI355X.

Headline workload (the largest single-GPU configuration): the synthetic
48-mode cubic+quartic field (19,600 cubic and 42,212 quartic constants with
frequency-scaled log-uniform magnitudes; Ntot 2, caps 6; eps1 1.0, Nstate
128, PT2 with eps2 0.2), generated with a fixed seed (vhci_amd/synth.py).
Its basis grows 1225 -> 1,003,305 -> 1,548,210 -> 1,598,434 states with
about 1.9e9 stored entries, and PT2 runs over 3.2 M external states.  One
step = one whole solve: heat-bath iterations (select -> build H ->
diagonalize) until the basis grows by at most 1 %, then Epstein-Nesbet PT2.
Inputs are generated on the host before the timed region; each step
creates its device context (tables and initial basis upload) inside it.

Both precision forms run: FP64 and mixed (FP32 H/SpMM with FP64
refinement, FP64 PT2 denominators and Kahan sums).  The headline value is
the mixed form's solves/s when every energy is within 0.01 cm^-1 of FP64,
the selected basis is identical and the mixed form is faster, else the FP64
figure.  The 15-mode workload (config 1) runs once per form beside it.

The serving entry runs a stream of requests drawn from the serving
workload's distribution (8-36 modes, eps1 0.5, 1 or 2, Nstate 20-200, eps2
0, 0.01 or 0.1; --serve requests): each request alone first (the solo FP64
rate), then Poisson arrivals at 1.2x that rate through a four-slot server,
once per precision form, with p50/p99 latency from submission to result and
a check that every result equals its solo run; then the capacity: all
requests at once through a four-slot FP64 server.

The reference-runs entry solves the inputs of the reference's own logged
runs (ClO2 and acetonitrile examples, the ethylene oxide run) and reports
each whole-program time beside the wall time its log prints.

The roofline entry is for the dominant kernel of the headline solve, the
FP64 eigensolver's block SpMM (k_spmm_lane), every launch of the timed
solves bracketed by HIP events on the context's stream; its HBM bytes per
launch come from two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over
one solve, scaled by the same counters' reading of a 1 GiB copy
(--no-traffic skips them).  The CPU baseline is the oracle's
pair-evaluation rate on a bounded sample of the headline run's states,
extrapolated to the whole solve with the run's exact pair counts.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import vhci_amd  # noqa: E402
from vhci_amd import synth  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "solves/s + p50/p99 latency, mixed vs FP64 at <=0.01 cm^-1 error; SpMV HBM GB/s"


def percentile(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else None


def run_form(problem, form, steps, warmup, bs):
    """Warmup solves, then `steps` timed solves.  The solver's bs-column
    SpMM launches of every timed solve are bracketed by HIP events on the
    context stream; returns (latencies, (last ctx, last result), spmm)."""
    for _ in range(warmup):
        ctx = vhci_amd.Context(problem, form)
        ctx.run()
        ctx.close()
    lat = []
    last = None
    sp = [0.0, 0.0, 0]
    gm = [0.0, 0.0, 0]
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx = vhci_amd.Context(problem, form)
        ctx.spmm_profile(bs)
        r = ctx.run()
        lat.append(time.perf_counter() - t0)
        ms, by, nl = ctx.spmm_profile_read()
        gms, gfl, gn = ctx.gemm_profile_read()
        ctx.spmm_profile(0)
        sp = [sp[0] + ms, sp[1] + by, sp[2] + nl]
        gm = [gm[0] + gms, gm[1] + gfl, gm[2] + gn]
        if last is not None:
            last[0].close()
        last = (ctx, r)
        print("[bench] form=%s solve %.3f s nvar=%d ns2=%d" % (form, lat[-1], r["nvar"], r["ns2"]),
              file=sys.stderr, flush=True)
    return lat, last, {"ms": sp[0], "bytes": sp[1], "launches": sp[2], "gemm_ms": gm[0], "gemm_flops": gm[1],
                       "gemm_calls": gm[2], "phases": last[0].phase_stats()}


FP64_PEAK_TFLOPS = 78.6


def phase_rooflines(ph, sp):
    """Algorithmic-bytes bandwidth of the H build, the PT2 selection and the
    PT2 sum of the last timed solve (wall times of the phases, syncs
    included), and the FP64 rate of the Rayleigh-Ritz GEMMs (the projection
    V^T W and the rotation X = V S, rocBLAS MFMA kernels)."""
    K = ph["key_bytes"]
    n, nnz, P = ph["n"], ph["nnz"], ph["patterns"]
    sv = 8.0  # FP64 value, or FP32 hi + FP32 lo
    probes = n * P  # every variational state's connections (PT2 probes them all)
    # an incremental build probes only the new rows and reads the carried entries once more
    build_probes = ph.get("probed_rows", n) * P
    hb = build_probes * (K + 4) + nnz * (sv + 4) + ph.get("carried", 0) * (sv + 4) + 8 * (n + 1) + n * K
    sel = n * (K + 8) + ph["hb_candidates"] * (K + 4) + ph["hb_unique"] * (2 * K + 4)
    ns = ph["nstate"]
    pt2 = n * (K + 8 * ns) + probes * (K + 4) + ph["pairs"] * 8 * ns * 2 + ph["ns2"] * (2 * K + 4 + 8 * ns)
    # The formulas count every pattern of every row as a (K+4)-byte probe.
    # Most patterns stop at the row's quanta check, the Bloom filter or the
    # 32-bit tag compare on cached data, so for the H build and the PT2 sum
    # the formula bytes per second exceed HBM bandwidth and are reported as
    # such (formula_GBs), with the pattern checks per second beside them;
    # the counted fabric traffic of these kernels is in the profiles.
    note = "formula bytes / time (every pattern counted as a probe); not HBM traffic"
    out = {"h_build": {"n": n, "stored_entries": nnz, "probed_rows": ph.get("probed_rows", n),
                       "carried_entries": ph.get("carried", 0), "pattern_checks": build_probes,
                       "checks_per_s": build_probes / ph["build_ms"] * 1e3 if ph["build_ms"] > 0 else None,
                       "formula_bytes": hb, "ms": ph["build_ms"],
                       "formula_GBs": hb / ph["build_ms"] / 1e6 if ph["build_ms"] > 0 else None, "note": note},
           "pt2_select": {"candidates": ph["hb_candidates"], "unique": ph["hb_unique"], "bytes": sel,
                          "ms": ph["hb_ms"], "GBs": sel / ph["hb_ms"] / 1e6 if ph["hb_ms"] > 0 else None},
           "pt2_sum": {"external": ph["ns2"], "pairs": ph["pairs"], "pattern_checks": probes,
                       "checks_per_s": probes / ph["pt2_sum_ms"] * 1e3 if ph["pt2_sum_ms"] > 0 else None,
                       "pairs_per_s": ph["pairs"] / ph["pt2_sum_ms"] * 1e3 if ph["pt2_sum_ms"] > 0 else None,
                       "formula_bytes": pt2, "ms": ph["pt2_sum_ms"],
                       "formula_GBs": pt2 / ph["pt2_sum_ms"] / 1e6 if ph["pt2_sum_ms"] > 0 else None, "note": note}}
    if sp["gemm_ms"] > 0:
        tf = sp["gemm_flops"] / sp["gemm_ms"] / 1e9
        out["ritz_gemm"] = {"calls": sp["gemm_calls"], "ms": sp["gemm_ms"], "TFLOPs": tf,
                            "frac_of_fp64_peak": tf / FP64_PEAK_TFLOPS}
    return out


def spmm_kernel_name(bs, mode, lanes=64):
    """The block SpMM kernel as profilers name it (lanes per row: 64, or 32
    for matrices whose rows average fewer than 128 entries)."""
    return "k_spmm_lane<%d, %d, %d>" % (bs, mode, lanes)


def traffic_pass(counters, kernel, workload, timeout_s=420):
    """One rocprofv3 counter pass (a single FP64 solve of this workload plus
    the 1 GiB calibration copy) in a child process; returns, per counter, the
    mean value per dispatch of `kernel` and of the copy kernel k_copy16, and
    the number of `kernel` dispatches."""
    import csv
    import glob
    import signal
    import subprocess
    import tempfile

    d = tempfile.mkdtemp(prefix="vhci_pmc_")
    cmd = ["rocprofv3", "--pmc"] + counters + ["--kernel-include-regex", "k_spmm_lane|k_copy16",
                                               "--output-format", "csv", "-d", d, "-o", "pmc", "--",
                                               sys.executable, "-u", os.path.abspath(__file__), "--steps", "1",
                                               "--warmup", "0", "--forms", "fp64", "--no-cpu-baseline",
                                               "--serve", "0", "--no-reference-runs", "--no-traffic",
                                               "--no-secondary", "--workload", workload]
    print("[bench] counter pass: %s" % " ".join(cmd), file=sys.stderr, flush=True)
    log = os.path.join(d, "pass.log")
    with open(log, "w") as lf:
        p = subprocess.Popen(cmd, stdout=lf, stderr=subprocess.STDOUT, start_new_session=True)
        try:
            rc = p.wait(timeout=timeout_s)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            raise RuntimeError("counter pass %s timed out (log %s)" % (counters, log))
    print("[bench] counter pass %s ended with status %d" % (counters, rc), file=sys.stderr, flush=True)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError("counter pass %s wrote no counter file (exit %d)" % (counters, rc))
    vals = {c: {"kernel": [], "copy": []} for c in counters}
    with open(files[0]) as f:
        for r in csv.DictReader(f):
            c = r["Counter_Name"]
            if c not in vals:
                continue
            name = r["Kernel_Name"]
            if kernel in name:
                vals[c]["kernel"].append(float(r["Counter_Value"]))
            elif "k_copy16" in name:
                vals[c]["copy"].append(float(r["Counter_Value"]))
    out = {}
    nk = 0
    for c, v in vals.items():
        if not v["kernel"] or not v["copy"]:
            raise RuntimeError("counter pass %s: kernel or copy dispatches missing" % c)
        out[c] = (sum(v["kernel"]) / len(v["kernel"]), sum(v["copy"]) / len(v["copy"]))
        nk = len(v["kernel"])
    return out, nk


def measure_traffic(kernel, workload, request_sizes=False):
    """HBM bytes per launch of `kernel`.  traffic: a FETCH_SIZE pass and a
    WRITE_SIZE pass (KiB per dispatch), each scaled by what the same counter
    reads for the 1 GiB copy (1 GiB fetched, 1 GiB written).  FETCH_SIZE
    counts a 128-byte read request as 64 bytes on this GPU (the copy reads
    half its size), while 32- and 64-byte requests count in full, so the
    copy-scaled figure overstates a kernel whose reads are 64-byte gathers;
    with request_sizes a third pass counts the read requests by size (32,
    64, 128 bytes) and gives the read bytes directly."""
    f, nk = traffic_pass(["FETCH_SIZE"], kernel, workload)
    w, _ = traffic_pass(["WRITE_SIZE"], kernel, workload)
    gib_kib = float(1 << 20)
    fk, fc = f["FETCH_SIZE"]
    wk, wc = w["WRITE_SIZE"]
    ff, fw = gib_kib / fc, gib_kib / wc
    traffic = 1024.0 * (fk * ff + wk * fw)
    out = {"fetch_kib": fk, "write_kib": wk, "dispatches": nk, "copy_fetch_kib": fc, "copy_write_kib": wc,
           "fetch_scale": ff, "write_scale": fw}
    if request_sizes:
        sizes = ["TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum"]
        r, _ = traffic_pass(sizes, kernel, workload)
        read_bytes = 32.0 * r[sizes[0]][0] + 64.0 * r[sizes[1]][0] + 128.0 * r[sizes[2]][0]
        out.update({"read_requests_32_64_128_all": [r[c][0] for c in sizes],
                    "read_bytes_by_request_size": read_bytes,
                    "copy_read_bytes_by_request_size": 32.0 * r[sizes[0]][1] + 64.0 * r[sizes[1]][1] +
                    128.0 * r[sizes[2]][1],
                    "traffic_by_request_size": read_bytes + 1024.0 * wk * fw})
    return traffic, out


def cpu_baseline(ctx, r, problem_text, threads):
    """Oracle pair rates on a bounded sample of this run's states,
    extrapolated with the run's exact pair counts (all-pairs H per
    iteration, |S2| x N PT2 gather)."""
    import tempfile

    from oracle import pyoracle

    pyoracle.set_threads(threads)
    with tempfile.NamedTemporaryFile("w", suffix=".inp", delete=False) as f:
        f.write(problem_text)
        path = f.name
    try:
        o = pyoracle.OProblem(path)
    finally:
        os.unlink(path)
    q = ctx.basis()
    ns2 = r["ns2"]
    # bounded sample, about 20 s of the oracle on the host cores: 3e8 pairs
    # of the all-pairs H loop and 1e9 of the PT2 gather
    nA = min(ns2, max(200, int(1e9 / max(1, q.shape[0]))))
    qA = ctx.new_states(ns2)[:nA] if nA else np.zeros((0, q.shape[1]), dtype=np.int32)
    n = q.shape[0]
    rows = min(n, max(1, int(3e8 / n)))
    nopt = min(o.neig, n)
    rate_h, rate_p, raw = o.sample_rates(q, rows, qA, nopt)
    sizes = ctx.iteration_sizes()
    pairs_h = sum(k * (k + 1) / 2.0 for k in sizes[1:]) if len(sizes) > 1 else sizes[0] * (sizes[0] + 1) / 2.0
    pairs_p = float(ns2) * n
    t = pairs_h / rate_h + (pairs_p / rate_p if nA else 0.0)
    sample_s = raw[1] + raw[3]
    return {
        "value": 1.0 / t,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": ("oracle all-pairs H rows 0..%d of the final %d-state basis (%.2e pairs, %.1f s) and PT2 gather "
                   "of %d external states (%.2e pairs, %.1f s); extrapolated to %.3e H pairs + %.3e PT2 pairs = "
                   "%.0f s per solve (eigensolver time not included)"
                   % (rows, n, raw[0], raw[1], nA, raw[2], raw[3], pairs_h, pairs_p, t)),
        "sample_seconds": sample_s,
    }


def cpu_fast_baseline(ctx, r, problem_text, threads, rows_h=4096, rows_p=2048):
    """The oracle's hash-based host form (the generate-and-lookup algorithm
    of the device engine on the host cores) timed on a bounded sample of this
    run's states and extrapolated with the run's own counts: H rows probed
    per build (new rows only when fewer than half are new, as the engine
    does), the eigensolver's product vectors at each iteration's basis size,
    and the PT2 pairs.  Heat-bath selection, Gram-Schmidt and Rayleigh-Ritz
    are not priced, which favours the CPU."""
    import tempfile

    from oracle import pyoracle

    pyoracle.set_threads(threads)
    with tempfile.NamedTemporaryFile("w", suffix=".inp", delete=False) as f:
        f.write(problem_text)
        path = f.name
    try:
        o = pyoracle.OProblem(path)
    finally:
        os.unlink(path)
    q = ctx.basis()
    n = q.shape[0]
    ns2 = int(r["ns2"])
    qA = ctx.new_states(ns2) if ns2 else np.zeros((0, q.shape[1]), dtype=np.int32)
    pairs_total = float(ctx.phase_stats()["pairs"]) if ns2 else 0.0
    nopt = min(o.neig, n)
    t0 = time.time()
    s = o.fast_sample(q, qA, rows_h, rows_p if ns2 else 0, nopt)
    sample_s = time.time() - t0
    (t_index, nh, t_h, ent_h, t_prod, t_index_a, npr, t_gen, npairs, t_acc, _chk, seen, t_ea) = s.tolist()
    sizes = ctx.iteration_sizes()
    probed = float(sizes[0])
    for a, b in zip(sizes[:-1], sizes[1:]):
        probed += (b - a) if 2 * (b - a) < b else b
    build = t_h / nh * probed + t_index * sum(sizes) / n
    # product vectors split evenly over the sparse iterations, each priced at
    # the sampled time per stored entry times the final H's entries per row
    # (diagonal included) times that iteration's basis size
    sparse = [k for k in sizes[1:]] or [sizes[0]]
    vectors = float(r["timings"].get("spmm_vectors", 0.0))
    nnz = float(ctx.phase_stats().get("nnz", 0.0))
    per_row = (nnz + n) / n if nnz > 0 else ent_h / nh
    eig = sum(vectors / 8.0 / len(sparse) * (t_prod / ent_h) * per_row * k for k in sparse)
    pt2 = 0.0
    if ns2 and npr > 0 and npairs > 0:
        pt2 = t_index_a + t_gen / npr * n + t_acc / npairs * pairs_total + t_ea / max(1.0, seen) * ns2
    t = build + eig + pt2
    return {
        "value": 1.0 / t,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "form": "hash-based (oracle fast form)",
        "sample": ("oracle fast form on this run's states: index of %d states (%.2f s); %d H rows (%.3e entries, "
                   "%.2f s; one 8-column product over them %.4f s); external index of %d states (%.2f s); PT2 pairs "
                   "of %d rows (%d pairs, generation %.2f s, accumulation %.2f s; E_a of the %d external states they "
                   "reach %.2f s); extrapolated to %d probed rows, %.0f product vectors, %.3e PT2 pairs and %d "
                   "external states: H builds %.0f s + products %.0f s + PT2 %.0f s = %.0f s per solve (heat-bath "
                   "selection, Gram-Schmidt and Rayleigh-Ritz not included)"
                   % (n, t_index, nh, ent_h, t_h, t_prod, ns2, t_index_a, npr, npairs, t_gen, t_acc, seen, t_ea,
                      probed, vectors, pairs_total, ns2, build, eig, pt2, t)),
        "phases_s": {"h_builds": build, "products": eig, "pt2": pt2},
        "sample_seconds": sample_s,
    }


# the serving workload's distribution; eps2 = 0 draws on fields of more than
# 20 modes become eps2 = 0.01 (synth.request_params: PT2 over every connected
# state of such a field holds 10^8 to 10^9 states)
SERVE_RANGES = dict(m_range=(8, 36), eps1_set=(0.5, 1.0, 2.0), nstate_range=(20, 200), eps2_set=(0.0, 0.01, 0.1),
                    eps2_zero_max_modes=20)


def serving(nreq, slots, load, seed=7, wide=False, alone_file=None, solo_only=False, forms=("fp64", "mixed"),
            saturated=True):
    """Serving stream of the serving workload's distribution (SERVE_RANGES;
    requests 0..nreq-1 of its seeded sequence): each request alone on a
    one-slot server sets the solo FP64 rate; then Poisson arrivals at `load`
    x that rate go through a `slots`-slot server, once per form.  Latency =
    submission to result.  alone_file: the solo results are read from it
    when it exists, else written to it (a long stream split over two runs)."""
    import threading

    texts = [synth.request_input(r, **SERVE_RANGES) for r in range(nreq)]
    alone = []
    if alone_file and os.path.exists(alone_file):
        saved = json.load(open(alone_file))
        if saved["requests"] < nreq:
            raise RuntimeError("%s holds %d solo results, %d needed" % (alone_file, saved["requests"], nreq))
        for a in saved["alone"][:nreq]:
            a["E"] = np.array(a["E"])
            a["checksum"] = int(a["checksum"])
            alone.append(a)
        print("[bench] serving: %d solo results read from %s" % (nreq, alone_file), file=sys.stderr, flush=True)
    else:
        srv = vhci_amd.Server(slots=1)
        for r, t in enumerate(texts):
            rid = srv.submit(t)
            srv.wait(rid)
            a = srv.result(rid)
            alone.append(a)
            print("[bench] serving request %d alone (M %d, eps1 %g, Nstate %d, eps2 %g): %d states, PT2 space %d, "
                  "%.2f s" % ((r,) + synth.request_params(r, **SERVE_RANGES)[:4] +
                              (a["nvar"], a["ns2"], a["t_end"] - a["t_start"])), file=sys.stderr, flush=True)
            if alone_file:
                # rewritten after every request, so a run cut short keeps what it finished
                keep = ("nvar", "ns2", "checksum", "t_start", "t_end", "peak_bytes")
                rows = [dict({k: x[k] for k in keep}, E=x["E"].tolist()) for x in alone]
                with open(alone_file + ".tmp", "w") as f:
                    json.dump({"requests": len(rows), "alone": rows}, f)
                os.replace(alone_file + ".tmp", alone_file)
        srv.close()
    if solo_only:
        return None
    service = [a["t_end"] - a["t_start"] for a in alone]
    rate = load * nreq / sum(service)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=nreq)
    out = {"requests": nreq, "slots": slots,
           "ranges": {k: (list(v) if isinstance(v, tuple) else v) for k, v in SERVE_RANGES.items()},
           "solo_fp64_solves_per_s": nreq / sum(service), "offered_rate_per_s": rate,
           "solo_p50_s": percentile(service, 50), "solo_p99_s": percentile(service, 99),
           "basis_sizes": [a["nvar"] for a in alone], "pt2_sizes": [a["ns2"] for a in alone]}
    for name, form in (("fp64", vhci_amd.FP64), ("mixed", vhci_amd.MIXED)):
        if name not in forms:
            continue
        srv = vhci_amd.Server(slots=slots)
        ids = [None] * nreq

        def feed():
            t0 = time.perf_counter()
            due = 0.0
            for i, t in enumerate(texts):
                due += gaps[i]
                dt = due - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(t, form=form)
                print("[bench] serving %s: request %d submitted" % (name, i), file=sys.stderr, flush=True)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res = []
        for k, rid in enumerate(ids):
            srv.wait(rid)
            res.append(srv.result(rid))
            print("[bench] serving %s: request %d done" % (name, k), file=sys.stderr, flush=True)
        st = srv.stats()
        srv.close()
        lat = [r["t_end"] - r["t_submit"] for r in res]
        span = max(r["t_end"] for r in res) - min(r["t_submit"] for r in res)
        same = sum(1 for a, r in zip(alone, res) if a["checksum"] == r["checksum"])
        err = max(float(np.max(np.abs(a["E"] - r["E"]))) for a, r in zip(alone, res))
        out[name] = {"solves_per_s": nreq / span, "p50_latency_s": percentile(lat, 50),
                     "p99_latency_s": percentile(lat, 99), "max_concurrent": st["max_running"],
                     "basis_identical_to_alone_fp64": same, "max_abs_err_vs_alone_fp64_cm": err}
        print("[bench] serving %s: %.2f solves/s p50 %.3f s p99 %.3f s" % (name, nreq / span, out[name]["p50_latency_s"],
                                                                       out[name]["p99_latency_s"]),
              file=sys.stderr, flush=True)
    if "mixed" in out and "fp64" in out:
        out["mixed_over_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    # capacity: every request submitted at once (FP64), served rate =
    # requests / makespan; the arrival stream above is paced below it
    out["saturated_fp64"] = {}
    for ns in (sorted({slots} | ({2 * slots} if wide else set())) if saturated else []):
        srv = vhci_amd.Server(slots=ns)
        ids = [srv.submit(t) for t in texts]
        res = []
        for k, rid in enumerate(ids):
            srv.wait(rid)
            res.append(srv.result(rid))
            print("[bench] serving fp64 saturated, %d slots: request %d done" % (ns, k), file=sys.stderr, flush=True)
        st = srv.stats()
        srv.close()
        lat = [r["t_end"] - r["t_submit"] for r in res]
        span = max(r["t_end"] for r in res) - min(r["t_submit"] for r in res)
        same = sum(1 for a, r in zip(alone, res) if a["checksum"] == r["checksum"])
        out["saturated_fp64"][str(ns)] = {"solves_per_s": nreq / span, "p50_latency_s": percentile(lat, 50),
                                          "p99_latency_s": percentile(lat, 99), "max_concurrent": st["max_running"],
                                          "basis_identical_to_alone_fp64": same}
        print("[bench] serving fp64 saturated, %d slots: %.2f solves/s" % (ns, nreq / span), file=sys.stderr,
              flush=True)
    return out


GOLDEN = os.path.join(ROOT, "tests", "golden")
# The reference's own logged runs (8 OpenMP threads; the CPU is not stated):
# (workload, input under tests/golden, header values for a force field or
# None, wall seconds the log prints, where the log says it).
REFERENCE_RUNS = [
    ("ClO2 conventional VCI (N = 3276, 20 states)", "clo2_vci.inp", None, 2.00,
     "examples/CLO2/VCI/output.log:38"),
    ("acetonitrile conventional VCI (N = 50,388, 70 states)", "acn_vci.inp", None, 2.60 * 60,
     "examples/Acetonitrile/VCI/output.log:38"),
    ("acetonitrile VHCI, eps1 1, 70 states", "acn_vhci.inp", None, 3.05 * 60,
     "examples/Acetonitrile/VHCI/output.log:66"),
    ("acetonitrile VHCI + PT2, eps1 1, eps2 0, 70 states", "acn_pt2.inp", None, 28.92 * 60,
     "examples/Acetonitrile/VHCI+PT2/output.log:74"),
    ("ethylene oxide VHCI + PT2, eps1 1, eps2 0.01, 200 states", "fields/ethylene_oxide.inp",
     ("1", "200", "1", "0.01", "2"), 6.80 * 3600, "the vim swap file of its run log, .ethylene_oxide_eps1-1.log.swp"),
]


def reference_runs():
    """The reference's logged runs solved here, whole program (input parse,
    context, heat-bath loop, PT2), FP64, against the wall time each log
    prints."""
    out = []
    for name, rel, header, ref_s, where in REFERENCE_RUNS:
        text = open(os.path.join(GOLDEN, rel)).read()
        if header:
            lines = text.split("\n")
            # header lines 2-6: HCI_Eps, Nstate, PT2, PT2_Eps, Max_quanta
            for i, v in enumerate(header):
                lines[1 + i] = lines[1 + i].split()[0] + " " + v
            text = "\n".join(lines)
        t0 = time.perf_counter()
        ctx = vhci_amd.Context(vhci_amd.Problem(text=text), vhci_amd.FP64)
        r = ctx.run()
        t = time.perf_counter() - t0
        ctx.close()
        out.append({"workload": name, "seconds": t, "reference_seconds": ref_s, "speedup": ref_s / t,
                    "reference_log": where, "nvar": r["nvar"], "ns2": r["ns2"]})
        print("[bench] %s: %.2f s (reference log %.0f s, %.0fx)" % (name, t, ref_s, ref_s / t), file=sys.stderr,
              flush=True)
    return out


def large_field_run():
    """The 48-mode naphthalene field at the large workload's parameters
    (eps1 1, Nstate 128, PT2 with eps2 0.2, total quanta 2): one FP64 solve
    with the block SpMM launches timed, for the behaviour at a few million
    states (the reference's published output of this run is a parity
    fixture of the tests)."""
    lines = open(os.path.join(GOLDEN, "fields", "naphthalene.inp")).read().split("\n")
    for i, v in enumerate(("1", "128", "1", "0.2", "2")):
        lines[1 + i] = lines[1 + i].split()[0] + " " + v
    t0 = time.perf_counter()
    ctx = vhci_amd.Context(vhci_amd.Problem(text="\n".join(lines)), vhci_amd.FP64)
    ctx.spmm_profile(8)
    r = ctx.run()
    t = time.perf_counter() - t0
    ms, by, nl = ctx.spmm_profile_read()
    ph = ctx.phase_stats()
    ctx.close()
    out = {"workload": "naphthalene (48 modes) VHCI + PT2, eps1 1, Nstate 128, eps2 0.2", "seconds": t,
           "nvar": r["nvar"], "ns2": r["ns2"], "stored_entries": ph["nnz"], "timings_ms": r["timings"],
           "spmm": {"kernel": spmm_kernel_name(8, 0, 32 if ph["nnz"] < 128 * ph["n"] else 64), "launches": nl, "avg_launch_ms": ms / max(nl, 1),
                    "GBs": by / ms / 1e6 if ms > 0 else None,
                    "frac_of_hbm": by / ms / 1e6 / HBM_PEAK_GBS if ms > 0 else None}}
    print("[bench] %s: %.2f s, %d states, %d PT2 states, %.3g entries, SpMM %.0f GB/s" %
          (out["workload"], t, r["nvar"], r["ns2"], ph["nnz"], out["spmm"]["GBs"] or 0), file=sys.stderr, flush=True)
    return out


WORKLOADS = {
    "c2": ("synthetic 48-mode cubic+quartic field, 19,600 cubic + 42,212 quartic FCs (log-uniform |f| with "
           "frequency-scaled ranges, |f| >= 0.5), Ntot 2, caps 6, eps1 1.0, Nstate 128, PT2 eps2 0.2 "
           "(seed 0x5648434902)", dict(modes=48, nstate=128, eps1=1.0, eps2=0.2)),
    "c1": ("synthetic 15-mode cubic+quartic field, 680 cubic + 600 quartic FCs, Ntot 2, caps 6, eps1 1.0, "
           "Nstate 100, PT2 eps2 0.01 (seed 0x5648434901)", dict(modes=15, nstate=100, eps1=1.0, eps2=0.01)),
}


def form_summary(lat, ctx, r, sp):
    return {"solves_per_s": 1.0 / float(np.mean(lat)), "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99),
            "nvar": r["nvar"], "ns2": r["ns2"], "hb_iterations": r["hb_iterations"], "timings_ms": r["timings"],
            "basis_sizes": ctx.iteration_sizes(), "E0": float(r["E"][0]),
            "solver_spmm": {"launches": sp["launches"], "ms": sp["ms"],
                            "GBs": sp["bytes"] / sp["ms"] / 1e6 if sp["ms"] > 0 else None}}


def compare_forms(r64, r32):
    err = float(np.max(np.abs(r64["E"] - r32["E"])))
    err_var = float(np.max(np.abs(r64["E_var"] - r32["E_var"])))
    same = r64["checksum"] == r32["checksum"] and r64["nvar"] == r32["nvar"] and r64["ns2"] == r32["ns2"]
    return {"max_abs_err_cm": err, "max_abs_err_var_cm": err_var, "basis_identical": same}


def secondary_run(name):
    """One solve per form of another workload (no warmup: the kernels are
    loaded by then), for continuity with earlier measurements."""
    text = synth.config_input(int(name[1:]))
    problem = vhci_amd.Problem(text=text)
    out = {"workload": WORKLOADS[name][0]}
    res = {}
    for fname, f in (("fp64", vhci_amd.FP64), ("mixed", vhci_amd.MIXED)):
        lat, (ctx, r), sp = run_form(problem, f, 1, 0, 8)
        out[fname] = form_summary(lat, ctx, r, sp)
        res[fname] = r
        ctx.close()
    out["mixed_vs_fp64"] = compare_forms(res["fp64"], res["mixed"])
    out["mixed_vs_fp64"]["speedup"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mixed-steps", type=int, default=1, help="timed solves of the mixed form (after the FP64 ones)")
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS), help="headline workload")
    ap.add_argument("--forms", default="both", choices=["both", "fp64", "mixed"])
    ap.add_argument("--spmm-reps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--serve", type=int, default=8, help="requests in the serving stream (0: skip)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving stream (one JSON line)")
    ap.add_argument("--no-reference-runs", action="store_true",
                    help="skip the reference's logged runs (acetonitrile, ethylene oxide, ClO2)")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 15-mode workload's solves")
    ap.add_argument("--large-field", action="store_true",
                    help="also run the 48-mode naphthalene field at the large workload's parameters")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=1.2, help="offered load / solo FP64 rate")
    ap.add_argument("--serve-alone-file", default=None,
                    help="solo results of the serving stream: read when the file exists, else written")
    ap.add_argument("--serve-solo-only", action="store_true",
                    help="with --serve-only: only the solo pass (writes --serve-alone-file)")
    ap.add_argument("--serve-forms", default="fp64,mixed", help="forms of the arrival-stream passes")
    ap.add_argument("--serve-no-saturated", action="store_true", help="skip the capacity passes")
    ap.add_argument("--no-traffic", action="store_true", help="skip the counter passes (traffic is then null)")
    ap.add_argument("--request-sizes", action="store_true",
                    help="a third counter pass: the SpMM's read requests by size (32, 64, 128 bytes)")
    ap.add_argument("--large-field-only", action="store_true",
                    help="only the 48-mode naphthalene run at the large workload's parameters (one JSON line)")
    ap.add_argument("--profile-spmm", action="store_true",
                    help="only build the final H of the workload (FP64), then time --spmm-reps block SpMMs and "
                         "one 1 GiB calibration copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.large_field_only:
        print(json.dumps(large_field_run()), flush=True)
        return
    if args.serve_only:
        if args.serve_solo_only:
            if not args.serve_alone_file or os.path.exists(args.serve_alone_file):
                sys.exit("--serve-solo-only needs a --serve-alone-file that does not exist yet")
            serving(args.serve, 1, args.serve_load, alone_file=args.serve_alone_file, solo_only=True)
            print(json.dumps({"serving_solo": args.serve, "file": args.serve_alone_file}), flush=True)
            return
        print(json.dumps(serving(args.serve, args.serve_slots, args.serve_load, wide=True,
                                 alone_file=args.serve_alone_file, forms=args.serve_forms.split(","),
                                 saturated=not args.serve_no_saturated)), flush=True)
        return
    wname = args.workload
    cnum = int(wname[1:])
    text = synth.config_input(cnum)
    problem = vhci_amd.Problem(text=text)
    bs = int(os.environ.get("VHCI_KS_BLOCK", "8"))
    # rows of at least 128 entries on average take a whole wave (64 lanes):
    # the case of both workloads' final matrices
    kname = spmm_kernel_name(bs, 0, 64)
    traffic = None
    if not args.no_traffic and not args.profile_spmm:
        # the counter passes run before this process touches the GPU; a pass
        # that fails (no rocprofv3, a refused counter) leaves traffic null
        try:
            traffic = measure_traffic(kname, wname, args.request_sizes)
        except Exception as e:  # noqa: BLE001
            print("[bench] traffic not measured: %s" % e, file=sys.stderr, flush=True)
            traffic = None
    if args.profile_spmm:
        ctx = vhci_amd.Context(vhci_amd.Problem(text=synth.config_input(cnum, pt2=False)), vhci_amd.FP64)
        ctx.run()
        ms, by, nl = ctx.time_spmm(bs, args.spmm_reps)
        cms = ctx.copy_probe(1 << 30)
        print(json.dumps({"n": ctx.n, "b": bs, "ms_per_product": ms, "launches_per_product": nl,
                          "algorithmic_bytes_per_product": by, "GBs": by / ms / 1e6, "copy_bytes": 1 << 30,
                          "copy_ms": cms, "copy_GBs": 2.0 * (1 << 30) / cms / 1e6}), flush=True)
        return
    forms = {"fp64": vhci_amd.FP64, "mixed": vhci_amd.MIXED}
    if args.forms != "both":
        forms = {args.forms: forms[args.forms]}
    res = {}
    for name, f in forms.items():
        if name == "mixed" and "fp64" in forms:
            res[name] = run_form(problem, f, args.mixed_steps, 0, bs)
        else:
            res[name] = run_form(problem, f, args.steps, args.warmup, bs)

    desc, params = WORKLOADS[wname]
    out = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "data": "synthetic",
           "config": dict({"workload": desc}, **params)}
    forms_out = {name: form_summary(lat, ctx, r, sp) for name, (lat, (ctx, r), sp) in res.items()}
    out["forms"] = forms_out
    head = "mixed" if "mixed" in res else "fp64"
    if "mixed" in res and "fp64" in res:
        cmp_ = compare_forms(res["fp64"][1][1], res["mixed"][1][1])
        cmp_["speedup"] = forms_out["mixed"]["solves_per_s"] / forms_out["fp64"]["solves_per_s"]
        out["mixed_vs_fp64"] = cmp_
        if (not (cmp_["max_abs_err_cm"] <= 0.01 and cmp_["basis_identical"])
                or forms_out["mixed"]["solves_per_s"] < forms_out["fp64"]["solves_per_s"]):
            head = "fp64"
    lat, (ctx, r), _ = res[head]
    out["value"] = forms_out[head]["solves_per_s"]
    out["unit"] = "solves/s"
    out["ms_per_step"] = 1000.0 * float(np.mean(lat))
    out["dtype"] = "f64" if head == "fp64" else "f32+f64"
    out["headline_form"] = head
    out["p50_latency_s"] = forms_out[head]["p50_s"]
    out["p99_latency_s"] = forms_out[head]["p99_s"]

    # Roofline of the dominant kernel: the FP64 eigensolver's bs-column SpMM,
    # every launch of the timed solves bracketed by HIP events on the
    # context's stream.  Algorithmic bytes per launch (summed over the
    # launches, whose H grows with the basis): nnz*(8+4) + 8*(N+1) + 8*N +
    # 2*N*bs*8 (values and columns, row pointers, diagonal, X read once and
    # Y written once).
    ctx64 = res["fp64"][1][0] if "fp64" in res else ctx
    if "fp64" in res:
        sp = res["fp64"][2]
        ach = sp["bytes"] / sp["ms"] / 1e6
        per = sp["bytes"] / sp["launches"]
        out["roofline"] = {"kernel": "%s (FP64 eigensolver block SpMM, %d vectors, CSR)" % (kname, bs),
                           "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": sp["ms"] / sp["launches"],
                           "launches": sp["launches"], "algorithmic_bytes_per_launch": per}
        out["phases"] = phase_rooflines(sp["phases"], sp)
        if traffic is not None:
            t, detail = traffic
            out["roofline"]["traffic"] = t
            detail["traffic_over_algorithmic"] = t / per
            if "traffic_by_request_size" in detail:
                detail["traffic_by_request_size_over_algorithmic"] = detail["traffic_by_request_size"] / per
            out["roofline"]["traffic_detail"] = detail
    # single-vector SpMV on the final H, and a plain 1 GiB device copy
    ms1, b1, _ = ctx64.time_spmm(1, args.spmm_reps)
    out["spmv"] = {"kernel": "k_spmv<0>", "ms": ms1, "bytes": b1, "GBs": b1 / ms1 / 1e6,
                   "frac_of_hbm": b1 / ms1 / 1e6 / HBM_PEAK_GBS, "n": ctx64.n}
    if "mixed" in res:
        ms32, b32, _ = res["mixed"][1][0].time_spmm(1, args.spmm_reps, fp32=True)
        out["spmv_fp32"] = {"kernel": "k_spmv<1>", "ms": ms32, "bytes": b32, "GBs": b32 / ms32 / 1e6}
    cms = ctx64.copy_probe(1 << 30)
    out["hbm_copy"] = {"kernel": "k_copy16", "bytes_moved": 2 * (1 << 30), "ms": cms,
                       "GBs": 2.0 * (1 << 30) / cms / 1e6}
    if not args.no_cpu_baseline:
        rr = res["fp64"][1][1] if "fp64" in res else r
        out["cpu_baseline"] = cpu_baseline(ctx64, rr, text, args.cpu_threads)
        out["cpu_baseline_fast"] = cpu_fast_baseline(ctx64, rr, text, args.cpu_threads)
        out["cpu_baseline_fast"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline_fast"]["value"]
    for name, (_, (c, _r), _s) in res.items():
        c.close()
    if not args.no_secondary and wname != "c1":
        out["secondary_c1"] = secondary_run("c1")
    if args.serve > 0:
        out["serving"] = serving(args.serve, args.serve_slots, args.serve_load)
    if not args.no_reference_runs:
        out["reference_runs"] = reference_runs()
    if args.large_field:
        out["large_field"] = large_field_run()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
