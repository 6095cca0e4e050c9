"""Benchmark: Gbutterfly/s and HBM roofline fraction of the MI355X NTT.

Headline step: one negacyclic BabyBear transform pair (forward then inverse)
of one n = 2^24 polynomial with uint32 coefficients resident in HBM (the
"single large negacyclic NTT" workload).  Butterflies per transform are
(n/2)*log2(n); algorithmic bytes per transform are 2*n*4 (one compulsory
read and write; twiddle tables are device-resident and not counted).

Extra lines in "extras": n = 2^12 cyclic latency through the uint64 ABI,
n = 2^20 single and batched, the 24-limb RNS batch, the 512-polynomial
uneven batch against one-polynomial-per-launch, and the uint64 ABI rate.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "ntt-tensorcore_amd"))
sys.path.insert(0, REPO)

import ntt_amd as nt  # noqa: E402
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
BB = W.BABYBEAR


def log(msg):
    print(msg, file=sys.stderr, flush=True)


RAMP_S = 0.3


def timed(fn, steps, warmup, stream):
    # Bring the GPU to steady clocks first (the step for RAMP_S seconds),
    # then the W counted warmup steps, then exactly K timed steps.
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < RAMP_S:
        fn()
        stream.sync()
    for _ in range(warmup):
        fn()
    stream.sync()
    e0, e1 = nt.Event(), nt.Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    ms = e0.elapsed_ms(e1)
    e0.destroy()
    e1.destroy()
    return ms


def headline(args, ctx, stream):
    log_n = args.log_n
    n = 1 << log_n
    x = W.random_poly(n, BB, 24, dtype=np.uint32)
    buf = nt.DeviceBuffer.from_array(x)
    fwd = nt.FLAG_NEGACYCLIC
    inv = nt.FLAG_NEGACYCLIC | nt.FLAG_INVERSE

    if args.force_split:
        split = [int(v) for v in args.force_split.split(",")]

        def step():
            ctx.transform_split(buf.ptr, BB, log_n, fwd, split, args.force_log_e, stream, mfma=args.force_mfma)
            ctx.transform_split(buf.ptr, BB, log_n, inv, split, args.force_log_e, stream, mfma=args.force_mfma)
        plan = {"split": split, "log_e": args.force_log_e, "mfma_mask": args.force_mfma, "forced": True}
    else:
        def step():
            ctx.transform(buf.ptr, BB, log_n, fwd, stream)
            ctx.transform(buf.ptr, BB, log_n, inv, stream)
        plan = nt.split_for(BB, log_n, n)

    ms = timed(step, args.steps, args.warmup, stream)
    stream.sync()
    ok = np.array_equal(buf.download(np.uint32, n), x)
    buf.free()
    _require(ok, "headline: the round trip after the timed region differs from the input")
    ms_step = ms / args.steps
    bfly = 2 * W.butterflies(log_n)
    t_transform_s = ms_step / 2 / 1e3
    # The dominant kernel is the (pipelined) pass kernel: a transform is P
    # back-to-back launches of it, each moving one pass over the coefficients.  One launch
    # processes 1/P of a transform's algorithmic bytes (2*n*4), so bytes per
    # launch / average launch time = 2*n*4 / transform time.
    npass = len(plan["split"])
    launch_s = t_transform_s / npass
    achieved = (2 * n * 4 / npass) / launch_s / 1e9
    traffic = None
    tf = os.path.join(REPO, "profiles", f"pipe_kernel_negacyclic_bb_2e{log_n}_traffic.json")
    if os.path.exists(tf):
        t = json.load(open(tf))
        if t.get("log_n") == log_n and t.get("split") == plan["split"] and t.get("log_e") == plan["log_e"] \
                and t.get("mfma_mask") == plan["mfma_mask"]:
            traffic = t["hbm_bytes_per_transform"]
    return {
        "plan": plan,
        "ms_per_step": ms_step,
        "value": bfly * args.steps / (ms / 1e3) / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "per": f"one pass-kernel launch: 2*n*4/{npass} algorithmic bytes over its average duration "
                            f"({npass} launches per transform); traffic = measured HBM bytes per transform",
                     "launches_per_transform": npass, "launch_us": launch_s * 1e6,
                     "transform_us": t_transform_s * 1e6},
    }


def cpu_baseline(args):
    from oracle import pyoracle as orc
    n = 1 << args.log_n
    t = orc.time_negacyclic_roundtrip(n, BB, 24, 1)
    return {"value": 2 * W.butterflies(args.log_n) / t / 1e9, "unit": "Gbutterfly/s", "cores": 1,
            "kind": "port",
            "sample": f"1 step: negacyclic forward+inverse round trip, n=2^{args.log_n}, oracle ntt_host "
                      f"restatement (radix-2 DIT, __int128 %), one thread (ntt_host is serial), {t:.2f} s",
            **host_info()}


def extra_latency_4096(ctx, stream):
    """cfg1: n=2^12 cyclic forward + inverse through the uint64 ABI, one call pair per step."""
    n = 4096
    cfg = nt.Config(n, BB)
    cfg.reserve(stream)
    buf = nt.DeviceBuffer.from_array(W.random_poly(n, BB, 42))
    lat = []
    for i in range(60):
        t0 = time.perf_counter()
        cfg.forward(buf, stream=stream)
        cfg.inverse(buf, stream=stream)
        stream.sync()
        if i >= 10:
            lat.append((time.perf_counter() - t0) * 1e6)
    ms = timed(lambda: (cfg.forward(buf, stream=stream), cfg.inverse(buf, stream=stream)), 200, 10, stream)
    cfg.cleanup()
    return {"p50_us_host_roundtrip": float(np.percentile(lat, 50)), "device_us_per_pair": ms * 1e3 / 200}


# The reference's only published measurement of this path: ntt_forward_optimized,
# forward cyclic BabyBear, one uint64_t polynomial per call, median us on an RTX
# 5060 (reference README.md:120-126).  Context only: other hardware.
REFERENCE_FWD_US = {12: 15.01, 14: 27.17, 16: 62.43, 17: 104.29, 18: 257.47, 19: 491.14, 20: 963.65}


def extra_reference_table(stream):
    """The reference's published workload on this library's drop-in uint64_t
    ABI (ntt_c_forward, variant optimized): one forward cyclic transform per
    call; device time per call from HIP events over 50 calls, median of 5."""
    out = {}
    for log_n, ref_us in REFERENCE_FWD_US.items():
        n = 1 << log_n
        cfg = nt.Config(n, BB)
        cfg.reserve(stream)
        buf = nt.DeviceBuffer.from_array(W.random_poly(n, BB, 42))
        step = lambda: cfg.forward(buf, stream=stream)  # noqa: E731
        reps = [timed(step, 50, 5, stream) * 1e3 / 50 for _ in range(5)]
        cfg.cleanup()
        buf.free()
        us = float(np.median(reps))
        out[f"2^{log_n}"] = {"us": us, "Gbfly_s": W.butterflies(log_n) / us / 1e3,
                             "reference_rtx5060_us": ref_us}
    return out


def extra_polymul(stream):
    """Negacyclic product a*b mod (x^n+1, BabyBear) on the uint64_t ABI
    (ntt_c_negacyclic_multiply: two forward transforms and an inverse whose
    first pass loads the pointwise product), against the same pipeline with the
    product as its own kernel (forward, forward, ntt_c_poly_multiply, inverse).
    Calls repeat on the same buffers (a and b stay residues mod q)."""
    out = {}
    for log_n in (12, 16, 20, 24):
        n = 1 << log_n
        cfg = nt.Config(n, BB)
        cfg.reserve(stream)
        a = nt.DeviceBuffer.from_array(W.random_poly(n, BB, 1))
        b = nt.DeviceBuffer.from_array(W.random_poly(n, BB, 2))
        r = nt.DeviceBuffer(n * 8)
        fused = lambda: cfg.negacyclic_multiply(r, a, b, stream=stream)  # noqa: E731
        split = lambda: (cfg.forward_negacyclic(a, stream), cfg.forward_negacyclic(b, stream),  # noqa: E731
                         cfg.poly_multiply(r, a, b, stream=stream), cfg.inverse_negacyclic(r, stream))
        reps = 20 if log_n < 24 else 5
        t_f = float(np.median([timed(fused, reps, 2, stream) * 1e3 / reps for _ in range(3)]))
        t_s = float(np.median([timed(split, reps, 2, stream) * 1e3 / reps for _ in range(3)]))
        for x in (a, b, r):
            x.free()
        cfg.cleanup()
        out[f"2^{log_n}"] = {"fused_us": t_f, "separate_pointwise_us": t_s, "speedup": t_s / t_f}
    return out


def extra_2_20(ctx, stream):
    out = {}
    for batch in (1, 16):
        n = 1 << 20
        buf = nt.DeviceBuffer(n * 4 * batch)
        for b in range(batch):
            buf.upload(W.random_poly(n, BB, 100 + b, dtype=np.uint32), b * n * 4)
        descs = [(buf.ptr + b * n * 4, BB, 20, nt.FLAG_NEGACYCLIC) for b in range(batch)]
        bt = nt.Batch(ctx, descs)
        ms = timed(lambda: bt.launch(stream), 50, 5, stream)
        t = ms / 50 / 1e3
        xs = [W.random_poly(n, BB, 100 + b) for b in range(batch)]
        expect, _ = _oracle().transform_batch(xs, [BB] * batch, [True] * batch, [False] * batch)

        def check(form):
            # one more launch on fresh inputs, checked against the oracle
            for b, x in enumerate(xs):
                buf.upload(x.astype(np.uint32), b * n * 4)
            bt.launch(stream)
            bt.check(stream)
            ok = all(np.array_equal(buf.download(np.uint32, n, b * n * 4).astype(np.uint64), e)
                     for b, e in enumerate(expect))
            _require(ok, f"negacyclic_2_20 batch {batch} ({form}): output differs from the oracle")

        check("per-phase launches")
        out[f"batch{batch}"] = {"Gbfly_s": batch * W.butterflies(20) / t / 1e9,
                                "GB_s": batch * 2 * n * 4 / t / 1e9, "us": t * 1e6, "checked": "oracle"}
        if batch > 1:
            out[f"batch{batch}"]["single_launch_queue_us"] = _queue_us(bt, stream)
            check("single work-queue launch")
        bt.destroy()
        buf.free()
    return out


def _queue_us(bt, stream):
    """The batch as one persistent work-queue launch (graph replay), timed,
    then checked against the oracle by the caller's next check; the batch is
    left in its default form (one launch per phase group)."""
    bt.set_queue(True)
    ms = timed(lambda: bt.launch_graph(stream), 20, 3, stream) / 20
    bt.check(stream)
    return ms * 1e3


def _queue_checked(bt, buf, offs, jobs, stream, name):
    """Times the single-launch form and checks its output against the oracle."""
    us = _queue_us(bt, stream)
    _, ok = _check_batch(buf, offs, jobs, lambda: bt.launch_graph(stream), stream)
    bt.check(stream)
    _require(ok, f"{name}: single-launch work-queue output differs from the oracle")
    bt.set_queue(False)
    return us


def _pack(jobs, dtype):
    esz = np.dtype(dtype).itemsize
    offs, tot = [], 0
    for j in jobs:
        offs.append(tot)
        tot += ((1 << j["log_n"]) * esz + 255) // 256 * 256
    buf = nt.DeviceBuffer(tot)
    for j, o in zip(jobs, offs):
        buf.upload(W.random_poly(1 << j["log_n"], j["q"], j["seed"], dtype=dtype), o)
    return buf, offs


def _flags(j, u64=False):
    return ((nt.FLAG_INVERSE if j["inverse"] else 0) | (nt.FLAG_NEGACYCLIC if j["nega"] else 0) |
            (nt.FLAG_U64 if u64 else 0))


def _oracle():
    """The CPU oracle, for the cpu_baseline legs: each extra that has a CPU
    baseline times the oracle on the same inputs, and the oracle's outputs
    double as the check of the GPU's."""
    from oracle import pyoracle as orc
    return orc


class CheckFailed(RuntimeError):
    pass


def _require(ok, msg):
    if not ok:
        raise CheckFailed(msg)


def _generator(q):
    """The smallest g >= 2 with g^((q-1)/p) != 1 for each p in {2, 3, 5, 7}
    dividing q - 1 (the generator rule of include/modular_arith.cuh:130-151)."""
    g = 2
    while any((q - 1) % p == 0 and pow(g, (q - 1) // p, q) == 1 for p in (2, 3, 5, 7)):
        g += 1
    return g


def _sampled_ok(got, x, q, nega):
    """A few outputs of a forward transform against the defining sums
    out[k] = sum_j x_j r_k^j (r_k = psi^(2k+1) or w^k), evaluated in numpy."""
    n = x.size
    g = _generator(q)
    for k in (0, 1, n // 3, n - 1):
        r = pow(pow(g, (q - 1) // (2 * n), q), 2 * k + 1, q) if nega else pow(pow(g, (q - 1) // n, q), k, q)
        p = np.ones(n, dtype=np.uint64)
        m, rm = 1, r
        while m < n:
            c = min(m, n - m)
            p[m:m + c] = (p[:c] * np.uint64(rm)) % np.uint64(q)
            rm = rm * rm % q
            m *= 2
        if int(got[k]) != int(((x.astype(np.uint64) * p) % np.uint64(q)).sum(dtype=np.uint64) % np.uint64(q)):
            return False
    return True


def _check_batch(buf, offs, jobs, run, stream):
    """Re-uploads the seeded inputs of a packed batch, runs it once and
    compares every polynomial with the oracle, whose transforms run one
    polynomial per worker on all host cores.  Returns (oracle seconds, ok)."""
    xs = [W.random_poly(1 << j["log_n"], j["q"], j["seed"]) for j in jobs]
    for x, o in zip(xs, offs):
        buf.upload(x.astype(np.uint32), o)
    run()
    stream.sync()
    expect, t = _oracle().transform_batch(xs, [j["q"] for j in jobs], [j["nega"] for j in jobs],
                                          [j["inverse"] for j in jobs])
    ok = all(np.array_equal(buf.download(np.uint32, 1 << j["log_n"], o).astype(np.uint64), e)
             for j, o, e in zip(jobs, offs, expect))
    return t, ok


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"nproc": nproc, "cpu_model": model}


def _cpu_pool(bfly, seconds, what):
    threads = _oracle().host_threads()
    return {"value": bfly / seconds / 1e9, "unit": "Gbutterfly/s", "cores": threads, "kind": "port",
            "sample": f"{what}: oracle ntt_host restatement, one polynomial per worker on {threads} threads, "
                      f"{seconds:.2f} s", **host_info()}


def extra_rns(ctx, stream):
    primes = W.rns_primes()
    jobs = [{"log_n": 16, "q": q, "inverse": False, "nega": True, "seed": 0x5EED ^ i} for i, q in enumerate(primes)]
    buf, offs = _pack(jobs, np.uint32)
    bt = nt.Batch(ctx, [(buf.ptr + o, j["q"], 16, _flags(j)) for j, o in zip(jobs, offs)])
    ms = timed(lambda: bt.launch_graph(stream), 50, 5, stream)
    t = ms / 50 / 1e3
    cpu, ok = _check_batch(buf, offs, jobs, lambda: bt.launch_graph(stream), stream)
    _require(ok, "rns_24x2_16: output differs from the oracle")
    queue_us = _queue_checked(bt, buf, offs, jobs, stream, "rns_24x2_16")
    bt.destroy()
    buf.free()
    n = 24 << 16
    bfly = 24 * W.butterflies(16)
    return {"limbs": 24, "log_n": 16, "us": t * 1e6, "single_launch_queue_us": queue_us, "Gbfly_s": bfly / t / 1e9,
            "GB_s": 2 * n * 4 / t / 1e9, "frac": 2 * n * 4 / t / 1e9 / HBM_PEAK_GBS, "checked": "oracle",
            "cpu_baseline": _cpu_pool(bfly, cpu, "24 limbs x 2^16 negacyclic forward")}


def extra_rns_kernels(ctx, stream):
    """RNS kernels around the transform on the 24 x 2^16 limb set: fast base
    conversion of 24 limbs to 24 other primes (ntt_c_bconv; 576 multiply-adds
    and 192 HBM bytes per coefficient), a 3-term per-limb multiply-accumulate
    (ntt_c_pmac, dnum = 3, accumulating: 32 HBM bytes per output residue), and
    that inner product fused into the 24 negacyclic inverse transforms it
    feeds (ntt_c_batch_prepare_ip) against ntt_c_pmac followed by the inverse
    batch."""
    n, L = 1 << 16, 24
    both = W.ntt_primes(1 << 31, 17, 2 * L)
    q, p = both[:L], both[L:]
    rng = np.random.default_rng(7)
    x = np.stack([rng.integers(0, v, size=n, dtype=np.uint64) for v in q]).astype(np.uint32)
    din = nt.DeviceBuffer.from_array(x)
    dout = nt.DeviceBuffer(L * n * 4)
    bc = nt.Bconv(q, p)
    t_b = timed(lambda: bc.run(dout, din, n, stream), 50, 5, stream) / 50 / 1e3
    dnum = 3
    a = nt.DeviceBuffer.from_array(np.concatenate([x] * dnum))
    b = nt.DeviceBuffer.from_array(np.concatenate([x[::-1]] * dnum))
    pm = nt.Pmac(q)
    t_p = timed(lambda: pm.run(dout, a, b, dnum, n, accumulate=True, stream=stream), 50, 5, stream) / 50 / 1e3
    # inner product -> inverse transform, fused and as two steps
    flags = nt.FLAG_INVERSE | nt.FLAG_NEGACYCLIC
    fused = nt.DeviceBuffer(L * n * 4)
    bf = nt.Batch(ctx, [(fused.ptr + l * n * 4, q[l], 16, flags) for l in range(L)],
                  ips=[(a.ptr + l * n * 4, b.ptr + l * n * 4, dnum, L * n) for l in range(L)])
    bs = nt.Batch(ctx, [(dout.ptr + l * n * 4, q[l], 16, flags) for l in range(L)])
    t_f = timed(lambda: bf.launch_graph(stream), 50, 5, stream) / 50 / 1e3

    def two_steps():
        pm.run(dout, a, b, dnum, n, accumulate=False, stream=stream)
        bs.launch_graph(stream)

    t_s = timed(two_steps, 50, 5, stream) / 50 / 1e3
    stream.sync()
    _require(np.array_equal(fused.download(np.uint32, L * n), dout.download(np.uint32, L * n)),
             "rns_kernels: fused inner product + inverse differs from pmac then inverse")
    for o in (bc, pm, bf, bs):
        o.destroy()
    for buf in (din, dout, a, b, fused):
        buf.free()
    bytes_b, bytes_p = n * 4 * (2 * L), n * L * 4 * (2 * dnum + 2)
    return {"bconv": {"L": L, "K": L, "n": n, "us": t_b * 1e6, "Gmac_s": n * L * L / t_b / 1e9,
                      "GB_s": bytes_b / t_b / 1e9, "frac": bytes_b / t_b / 1e9 / HBM_PEAK_GBS},
            "pmac": {"dnum": dnum, "L": L, "n": n, "us": t_p * 1e6, "GB_s": bytes_p / t_p / 1e9,
                     "frac": bytes_p / t_p / 1e9 / HBM_PEAK_GBS},
            "ip_inverse": {"dnum": dnum, "L": L, "n": n, "fused_us": t_f * 1e6, "pmac_then_inverse_us": t_s * 1e6,
                           "speedup": t_s / t_f, "checked": "fused == pmac then inverse (bit-exact)"}}


def extra_uneven(ctx, stream):
    jobs = W.uneven_batch()
    buf, offs = _pack(jobs, np.uint32)
    descs = [(buf.ptr + o, j["q"], j["log_n"], _flags(j)) for j, o in zip(jobs, offs)]
    bt = nt.Batch(ctx, descs)
    bt.launch_graph(stream)
    stream.sync()
    ms_b = timed(lambda: bt.launch_graph(stream), 20, 3, stream) / 20

    def per_poly_py():
        for d in descs:
            ctx.transform(d[0], d[1], d[2], d[3], stream)

    arr = nt.Context.desc_array(descs)

    def per_poly():  # the reference's per-polynomial loop, issued from native code
        ctx.transform_each(arr, stream)

    ms_p = timed(per_poly, 5, 1, stream) / 5
    ms_py = timed(per_poly_py, 3, 1, stream) / 3
    bfly = sum(W.butterflies(j["log_n"]) for j in jobs)
    byt = sum(2 * (1 << j["log_n"]) * 4 for j in jobs)
    info = bt.info()
    cpu, ok = _check_batch(buf, offs, jobs, lambda: bt.launch_graph(stream), stream)
    _require(ok, "uneven_512: batched output differs from the oracle")
    _, ok = _check_batch(buf, offs, jobs, per_poly, stream)
    _require(ok, "uneven_512: per-polynomial output differs from the oracle")
    _, ok = _check_batch(buf, offs, jobs, per_poly_py, stream)
    _require(ok, "uneven_512: per-polynomial (Python loop) output differs from the oracle")
    queue_us = _queue_checked(bt, buf, offs, jobs, stream, "uneven_512")
    bt.destroy()
    buf.free()
    return {"polys": len(jobs), "batched_us": ms_b * 1e3, "single_launch_queue_us": queue_us,
            "per_poly_us": ms_p * 1e3,
            "speedup": ms_p / ms_b, "per_poly_python_loop_us": ms_py * 1e3,
            "speedup_vs_python_loop": ms_py / ms_b, "batched_Gbfly_s": bfly / (ms_b / 1e3) / 1e9,
            "batched_GB_s": byt / (ms_b / 1e3) / 1e9, "launches": info["launches"],
            "per_poly_launches": sum(1 if j["log_n"] <= 14 else 2 for j in jobs), "checked": "oracle",
            "cpu_baseline": _cpu_pool(bfly, cpu, "512 uneven polynomials, log2 n 10..18, negacyclic")}


def extra_u64_abi(args, stream):
    log_n = args.log_n
    n = 1 << log_n
    cfg = nt.Config(n, BB)
    cfg.reserve(stream)
    buf = nt.DeviceBuffer.from_array(W.random_poly(n, BB, 24))
    ms = timed(lambda: (cfg.forward_negacyclic(buf, stream), cfg.inverse_negacyclic(buf, stream)), 10, 2, stream)
    stream.sync()
    x = W.random_poly(n, BB, 24)
    ok = np.array_equal(buf.download(np.uint64, n), x)
    _require(ok, "u64_abi: forward+inverse pairs do not round-trip")
    cfg.forward_negacyclic(buf, stream)
    stream.sync()
    ok = _sampled_ok(buf.download(np.uint64, n), x, BB, nega=True)
    _require(ok, "u64_abi: forward output differs from the defining sums")
    cfg.cleanup()
    buf.free()
    t = ms / 10 / 2 / 1e3
    return {"log_n": log_n, "Gbfly_s": W.butterflies(log_n) / t / 1e9, "GB_s": 2 * n * 8 / t / 1e9,
            "frac": 2 * n * 8 / t / 1e9 / HBM_PEAK_GBS, "checked": "round trip + sampled defining sums"}


def extra_klimb(stream):
    """K-limb big-prime path (include/ntt_arbitrary.cuh): forward + inverse over
    Montgomery-form limbs, n = 2^20 for the 48-bit (K=2) and Goldilocks (K=3)
    fields, n = 2^16 for BN254 (K=11).  Algorithmic bytes: 2 n 4K per transform.
    K = 2 and 3 run on the four-step pass kernels (ntt_c_arb_plan_*), timed
    beside the limb kernels (ntt_c_arb_forward / inverse) whose output they
    must equal; K = 11 has only the limb kernels."""
    import ntt_arb as A
    out = {}
    for K, log_n in ((2, 20), (3, 20), (11, 16)):
        n = 1 << log_n
        f = A.Field(A.FIELDS[K][0], K)
        limb = A.ArbNtt(f, log_n)
        fast = A.ArbPlan(limb) if K <= 3 else None
        rng = np.random.default_rng(K)
        x = rng.integers(0, 1 << 24, size=(n, K), dtype=np.uint32)
        x[:, -1] = rng.integers(0, int(f.p_limbs[-1]), size=n, dtype=np.uint32)
        buf = nt.DeviceBuffer.from_array(x)

        def run_of(t):
            return lambda: (t.forward(buf, stream), t.inverse(buf, stream))
        rec = {}
        for name, t in (("limb_kernels", limb), ("plan", fast)):
            if t is None:
                continue
            ms = timed(run_of(t), 10, 2, stream)
            stream.sync()
            _require(np.array_equal(buf.download(np.uint32, n * K).reshape(n, K), x), f"klimb K={K} {name} round trip")
            rec[name] = ms / 10 / 2 / 1e3
        if fast is not None:  # the plan's forward equals the limb kernels' forward, limb for limb
            other = nt.DeviceBuffer.from_array(x)
            fast.forward(buf, stream)
            limb.forward(other, stream)
            stream.sync()
            _require(np.array_equal(buf.download(np.uint32, n * K), other.download(np.uint32, n * K)),
                     f"klimb K={K} plan forward != limb kernels")
            other.free()
            fast.destroy()
        buf.free()
        t = rec.get("plan", rec["limb_kernels"])
        r = {"us": t * 1e6, "Gbfly_s": W.butterflies(log_n) / t / 1e9, "GB_s": 2 * n * 4 * K / t / 1e9,
             "kernels": "four-step pass kernels (plan)" if "plan" in rec else "limb kernels",
             "checked": "round trip" + (" + plan forward == limb kernels' forward" if "plan" in rec else "")}
        if "plan" in rec:
            r["limb_kernels_us"] = rec["limb_kernels"] * 1e6
            r["speedup_vs_limb_kernels"] = rec["limb_kernels"] / rec["plan"]
        out[f"K{K}_2^{log_n}"] = r
    return out


def extra_serving(ctx, qps_list=(100, 500, 1000, 2000, 4000, 8000), count=120):
    """cfg5: Poisson arrivals, 1..32 polys per request, log2 n in [12, 22],
    BabyBear (uint32) or a 62-bit prime (uint64); device-resident operands."""
    base = W.serving_requests(count, 1.0, seed=5)
    sizes = []
    for _, polys in base:
        sizes.append([(1 << p["log_n"]) * (8 if p["q"] > (1 << 31) else 4) for p in polys])
    total = sum((b + 255) // 256 * 256 for r in sizes for b in r)
    pool = nt.DeviceBuffer(total)
    nt.check(nt.lib().ntt_c_memset(pool.ptr, 1, total))  # 0x01.. < q for both prime kinds
    reqs, offs, off = [], [], 0
    for (_, polys), szs in zip(base, sizes):
        r, ro = [], []
        for p, b in zip(polys, szs):
            u64 = p["q"] > (1 << 31)
            flags = nt.FLAG_NEGACYCLIC | (nt.FLAG_INVERSE if p["inverse"] else 0) | (nt.FLAG_U64 if u64 else 0)
            r.append((pool.ptr + off, p["q"], p["log_n"], flags))
            ro.append(off)
            off += (b + 255) // 256 * 256
        reqs.append(r)
        offs.append(ro)
    bfly = sum(W.butterflies(p["log_n"]) for _, polys in base for p in polys)
    srv = nt.Server(ctx, hbm_budget_bytes=16 << 30, max_inflight=4, max_batch_polys=1024)
    srv.replay(reqs, [0.0] * len(reqs))  # builds every plan once (twiddles resident)
    out = {"requests": count, "polys": sum(len(r) for r in reqs), "butterflies": bfly, "points": []}
    for qps in qps_list:
        srv.reset_stats()
        arrivals = [t / qps for t, _ in base]
        lat = srv.replay(reqs, arrivals)
        st = srv.stats()
        out["points"].append({"qps": qps, "p50_us": float(np.percentile(lat, 50)),
                              "p99_us": float(np.percentile(lat, 99)), "Gbfly_s": bfly / st["wall_s"] / 1e9,
                              "batches": st["batches"], "max_concurrent_batches": st["max_concurrent_batches"]})
        log(f"[bench]   qps={qps}: p50 {out['points'][-1]['p50_us']:.0f} us p99 {out['points'][-1]['p99_us']:.0f} us")
    # check: seeded inputs for the requests of up to 2^21 coefficients, one
    # more replay of those (all arriving at once, so they share batches), and
    # every result against the oracle
    pick = [i for i, (_, polys) in enumerate(base) if sum(1 << p["log_n"] for p in polys) <= (1 << 21)]
    flat = [(p, o) for i in pick for p, o in zip(base[i][1], offs[i])]
    xs = [W.random_poly(1 << p["log_n"], p["q"], p["seed"]) for p, _ in flat]
    for (p, o), x in zip(flat, xs):
        pool.upload(x.astype(np.uint64 if p["q"] > (1 << 31) else np.uint32), o)
    srv.replay([reqs[i] for i in pick], [0.0] * len(pick))
    expect, t_cpu = _oracle().transform_batch(xs, [p["q"] for p, _ in flat], [True] * len(flat),
                                              [p["inverse"] for p, _ in flat])
    ok = all(np.array_equal(pool.download(np.uint64 if p["q"] > (1 << 31) else np.uint32, 1 << p["log_n"], o)
                            .astype(np.uint64), e) for (p, o), e in zip(flat, expect))
    _require(ok, "serving: a result differs from the oracle")
    out["checked"] = f"oracle, {len(pick)} requests ({len(flat)} polynomials) replayed together"
    out["cpu_baseline"] = _cpu_pool(sum(W.butterflies(p["log_n"]) for p, _ in flat), t_cpu,
                                    f"the {len(pick)} checked requests")
    srv.destroy()
    pool.free()
    return out


def tune_all(ctx, args):
    """Autotune every launch class the bench measures (choices keyed by size and
    elements in flight); returns the chosen splits next to the heuristic."""
    out = {}
    def rec(name, q, log_ns, elems):
        for r in nt.autotune(ctx, q, log_ns, elems, reps=4):
            out[f"{name}:2^{r['log_n']}"] = {k: r[k] for k in ("split", "log_e", "mfma_mask", "us", "heuristic_us")}
    rec("single", BB, sorted({args.log_n, *range(12, 24)}), 1)
    rec("batch_2^24", BB, [23, 20, 16], 1 << 24)
    rec("rns", BB, [16], 24 << 16)
    rec("uneven", BB, list(range(10, 19)), 1 << 24)
    rec("serve_bb", BB, list(range(12, 23)), 1 << 24)
    rec("serve_62", W.p62_primes()[0], list(range(12, 23)), 1 << 24)
    rec("single_62", W.p62_primes()[0], list(range(12, 23)), 1)
    rec("single_gold", 0xFFFFFFFF00000001, list(range(12, 23)), 1)  # Goldilocks (and K = 3 limb plans)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--retune", action="store_true", help="re-run the autotuner instead of loading its table")
    ap.add_argument("--force-split", default="", help="headline radix split, e.g. 12,12 (profiling other plans)")
    ap.add_argument("--force-log-e", type=int, default=14)
    ap.add_argument("--force-mfma", type=int, default=0, help="mask of passes on the matrix cores")
    ap.add_argument("--tune-out", default="", help="write the tuned table here")
    ap.add_argument("--extra", default="", help="run only this secondary measurement (profiling one kernel) "
                                                "and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the NTT path runs on a single MI355X "
              "(sharding across GPUs is not built)", file=sys.stderr)
        sys.exit(2)
    if nt.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    stream = nt.Stream()
    ctx = nt.Context(stream=stream)  # reserves the bench stream's workspace
    tuned = {}
    table = os.path.join(REPO, "ntt-tensorcore_amd", "tuned_mi355x.txt")
    if args.retune or not os.path.exists(table):
        log("[bench] autotune")
        tuned = tune_all(ctx, args)
        if args.tune_out:
            nt.autotune_save(args.tune_out)
    else:
        nt.autotune_load(table)
        tuned = {"loaded": os.path.relpath(table, REPO)}
    extra_fns = [("cyclic_4096", lambda: extra_latency_4096(ctx, stream)),
                 ("reference_table", lambda: extra_reference_table(stream)),
                 ("negacyclic_multiply", lambda: extra_polymul(stream)),
                 ("negacyclic_2_20", lambda: extra_2_20(ctx, stream)),
                 ("rns_24x2_16", lambda: extra_rns(ctx, stream)),
                 ("rns_kernels", lambda: extra_rns_kernels(ctx, stream)),
                 ("uneven_512", lambda: extra_uneven(ctx, stream)),
                 ("u64_abi", lambda: extra_u64_abi(args, stream)),
                 ("klimb", lambda: extra_klimb(stream)),
                 ("serving", lambda: extra_serving(ctx))]
    if args.extra:
        fns = dict(extra_fns)
        if args.extra not in fns:
            print(f"unknown --extra {args.extra}; one of {', '.join(fns)}", file=sys.stderr)
            sys.exit(2)
        print(json.dumps({"extra": args.extra, "result": fns[args.extra]()}), flush=True)
        ctx.destroy()
        stream.destroy()
        return
    log(f"[bench] headline: negacyclic n=2^{args.log_n} fwd+inv, {args.warmup} warmup + {args.steps} steps")
    h = headline(args, ctx, stream)
    log(f"[bench] headline done: {h['value']:.1f} Gbfly/s, {h['roofline']['achieved']:.0f} GB/s")
    extras = {}
    if not args.no_extras:
        for name, fn in extra_fns:
            log(f"[bench] extra {name}")
            extras[name] = fn()
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline (oracle, one thread)")
        cpu = cpu_baseline(args)
    out = {
        "metric": "Gbutterfly/s + %HBM roofline, batched vs per-poly launch; p99 latency at QPS",
        "value": h["value"], "unit": "Gbutterfly/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (splitmix64-seeded uniform coefficients in [0,q))",
        "config": {"workload": f"negacyclic NTT over BabyBear q=2013265921, n=2^{args.log_n}, forward then "
                               "inverse, one polynomial, uint32 coefficients resident in HBM",
                   "log_n": args.log_n, "q": BB, "transform": "negacyclic", "step": "forward+inverse"},
        "roofline": h["roofline"], "plan": h["plan"], "cpu_baseline": cpu, "extras": extras, "autotune": tuned,
    }
    print(json.dumps(out), flush=True)
    ctx.destroy()
    stream.destroy()


if __name__ == "__main__":
    try:
        main()
    except CheckFailed as e:
        print(f"[bench] CHECK FAILED: {e}", file=sys.stderr, flush=True)
        sys.exit(3)
