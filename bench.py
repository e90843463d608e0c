"""Benchmark of PDL's broadcast-loop path on one MI355X.

Workload ("broadcast + reductions at HBM scale"):
  x float32 (4096,4096,16) normal(0,1); y float32 (4096,16) -> y->dummy(1)
  z = x->xchg(0,1) * y->dummy(1)                    (transposed view * dummy dim)
  sumover(z), average(z), minimum(z), maximum(z)    (along dim 0)
  int32 (4096,4096) uniform[-2^20,2^20]: sumover, cumusumover   (bit-exact)
  int16 (8192,8192), 1% BAD (SHRT_MIN) + one all-bad row: sumover -> int32
One step = that whole chain; inputs resident in HBM, outputs preallocated.
value = algorithmic bytes of the step / step time (GB/s).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-matmult]
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
METRIC = "binop/sumover HBM GB/s (%peak); matmult TFLOP/s split-FP16 vs FP32-SIMT @err"


def algorithmic_bytes(nx, ny, nk, ni, nj, mi, mj):
    """distinct elements of every operand and output x size (dummy reuse counted once)"""
    f4, i4, i2 = 4, 4, 2
    binop = nx * ny * nk * f4 * 2 + nx * nk * f4
    red = nx * ny * nk * f4 + ny * nk * f4
    isum = ni * nj * i4 + nj * i4
    iscan = ni * nj * i4 * 2
    bsum = mi * mj * i2 + mj * i4
    return {"binop": binop, "sumover": red, "average": red, "minimum": red, "maximum": red,
            "i32_sumover": isum, "i32_cumusumover": iscan, "i16bad_sumover": bsum}


def measured_traffic(op):
    """(HBM bytes per launch, kernel) of this op from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy), or (None, None)"""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            d = json.load(f)
        return d["bytes_per_launch"].get(op), d.get("kernels", {}).get(op)
    except (OSError, ValueError, KeyError):
        return None, None


# ---------------------------------------------------------------- inputs
# Every workload's inputs are drawn on the device from PDL's own stream
# (srandom / random / grandom: one xoshiro256+ stream, the one PDL draws
# with autopthread off), written as the PDL expressions that make them:
#   uniform[-1,1]  random(type, dims) * 2 - 1
#   normal(0,1)    float(grandom(dims))   (double draws, converted once: a
#                  float grandom would round draws within 2^-25 of 1 to 1.0f
#                  and give +Inf, as PDL's does)

def uniform_pm1(C, *dims, dtype):
    return C.random(*dims, dtype=dtype) * 2 - 1


def normal_f32(C, *dims):
    return C.grandom(*dims).convert(C.L.F)


def make_int_inputs(C, ni, nj, mi, mj):
    """int32 (ni,nj) uniform in [-2^20, 2^20]: long(random(ni,nj) * 2097153) - 1048576;
    int16 (mi,mj) in [-32767, 32767]: short(long(random(mi,mj) * 65535) - 32767),
    then setbadif(random(mi,mj) < 0.01, and all of row 17) -- an all-bad row"""
    i32 = (C.random(ni, nj) * 2097153).convert(C.L.L) - 1048576
    v = ((C.random(mi, mj) * 65535).convert(C.L.L) - 32767).convert(C.L.S)
    mask = C.random(mi, mj) < 0.01
    C._run("assgn", [C.topdl(1), mask.slice(None, (17, 17, 1))])
    return i32, v.setbadif(mask)


def run_gpu(args):
    from pdl_amd import core as C

    C.init(0)
    NX, NY, NK = 4096, 4096, 16
    NI, NJ = 4096, 4096
    MI, MJ = 8192, 8192
    C.srandom(1)
    x = normal_f32(C, NX, NY, NK)                                 # PDL dims (4096,4096,16)
    y = normal_f32(C, NY, NK)                                     # (4096,16)
    xv = x.xchg(0, 1)
    yv = y.dummy(1)
    z = C.ndarray.new_dense(C.L.F, [NY, NX, NK])
    outs = {k: C.ndarray.new_dense(C.L.F, [NX, NK]) for k in ("sumover", "average", "minimum", "maximum")}
    i32, i16 = make_int_inputs(C, NI, NJ, MI, MJ)
    i32h, i16h = i32.numpy(), i16.numpy()
    assert i16.state & C.L.BADVAL and np.all(i16h[17] == -32768)
    is_out = C.ndarray.new_dense(C.L.L, [NJ])
    ic_out = C.ndarray.new_dense(C.L.L, [NI, NJ])
    ib_out = C.ndarray.new_dense(C.L.L, [MJ])
    ib_out.badflag(True)

    ops = [
        ("binop", lambda: C._run("mult", [xv, yv, z])),
        ("sumover", lambda: C._run("sumover", [z, outs["sumover"]])),
        ("average", lambda: C._run("average", [z, outs["average"]])),
        ("minimum", lambda: C._run("minimum", [z, outs["minimum"]])),
        ("maximum", lambda: C._run("maximum", [z, outs["maximum"]])),
        ("i32_sumover", lambda: C._run("sumover", [i32, is_out])),
        ("i32_cumusumover", lambda: C._run("cumusumover", [i32, ic_out])),
        ("i16bad_sumover", lambda: C._run("sumover", [i16, ib_out])),
    ]
    nbytes = algorithmic_bytes(NX, NY, NK, NI, NJ, MI, MJ)
    total_bytes = sum(nbytes.values())
    for _ in range(args.warmup):
        for _, f in ops:
            f()
    C.synchronize()
    # the timed steps: the ops back to back, events only around the whole
    t0 = time.perf_counter()
    C.synchronize()
    w0 = C.Event().record()
    for s in range(args.steps):
        for _, f in ops:
            f()
    w1 = C.Event().record()
    C.synchronize()
    t1 = time.perf_counter()
    step_ms_events = w0.elapsed_ms(w1) / args.steps
    step_ms_wall = (t1 - t0) * 1000 / args.steps
    # per-op breakdown (and the roofline's kernel time) from as many further
    # steps with an event pair around each op; the event markers between
    # kernels stall the queue a few microseconds each, so these steps are
    # not the timed ones
    evs = [[(C.Event(), C.Event()) for _ in ops] for _ in range(args.steps)]
    for s in range(args.steps):
        for k, (_, f) in enumerate(ops):
            evs[s][k][0].record()
            f()
            evs[s][k][1].record()
    C.synchronize()
    per_op = {}
    for k, (name, _) in enumerate(ops):
        ms = float(np.mean([evs[s][k][0].elapsed_ms(evs[s][k][1]) for s in range(args.steps)]))
        gbs = nbytes[name] / (ms * 1e-3) / 1e9
        per_op[name] = {"ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 3)}
    ms_per_step = step_ms_wall
    value = total_bytes / (ms_per_step * 1e-3) / 1e9
    dom = max(per_op, key=lambda k: per_op[k]["ms"])
    # parity spot check of this run's results (bit-exact parts)
    check = {}
    got = is_out.numpy()
    check["i32_sumover_exact"] = bool(np.array_equal(got, i32h.astype(np.int64).sum(axis=1).astype(np.int32)))
    got = ic_out.numpy()
    check["i32_cumusumover_exact"] = bool(np.array_equal(got, np.cumsum(i32h.astype(np.int64), axis=1).astype(np.int32)))
    good = i16h != -32768
    want = np.where(good, i16h.astype(np.int64), 0).sum(axis=1)
    want = np.where(good.any(axis=1), want, -2**31).astype(np.int32)
    check["i16bad_sumover_exact"] = bool(np.array_equal(ib_out.numpy(), want))
    # sampled k-planes of z (bit-exact float32 products) and of the four float
    # reductions (sums within the compensated bound, min/max exact)
    sums = {k: outs[k].numpy() for k in outs}
    z_ok, red_ok = True, True
    for k in (0, NK // 2, NK - 1):
        xk = x.slice(None, None, (k, k, 1)).numpy()[0]          # xk[j, i] = x(i, j, k)
        yk = y.slice(None, (k, k, 1)).numpy()[0]                # yk[i] = y(i, k)
        zk = z.slice(None, None, (k, k, 1)).numpy()[0]          # zk[j, i] = z(i, j, k) = x(j, i, k) * y(i, k)
        want = xk.T * yk[None, :]
        z_ok &= bool(np.array_equal(zk.view(np.uint32), want.view(np.uint32)))
        tr = want.astype(np.float64).sum(axis=1)
        ab = np.abs(want).astype(np.float64).sum(axis=1)
        for name, t, a in (("sumover", tr, ab), ("average", tr / NX, ab / NX)):
            bound = 2 * 2.0 ** -23 * np.abs(t) + 64 * 2.0 ** -46 * a + 1e-30
            red_ok &= bool(np.all(np.abs(sums[name][k].astype(np.float64) - t) <= bound))
        red_ok &= bool(np.array_equal(sums["minimum"][k], want.min(axis=1)))
        red_ok &= bool(np.array_equal(sums["maximum"][k], want.max(axis=1)))
    check["z_sampled_planes_exact"] = z_ok
    check["f32_reductions_sampled_planes"] = red_ok
    return {
        "value": value, "ms_per_step": ms_per_step, "ms_per_step_events": step_ms_events,
        "per_op": per_op, "dominant": dom, "dominant_bytes": nbytes[dom], "total_bytes": total_bytes,
        "check": check,
    }


def run_matmult(n=8192, nrows_check=16, reps=3):
    """Precision leg: float32 A,B (n,n) normal(0,1), and the same with rows of
    A scaled by exp(3*N(0,1)) (lognormal(0,3)); split-FP16 MFMA vs FP32 SIMT
    (FMA) vs opt-in FP8, each with TFLOP/s and relative Frobenius error vs
    the FP64 product on a sample of rows; PDL's own FP32 loop (the oracle)
    scored on the same rows.  Double (n/2,n/2) on the FP64 SIMT path."""
    from pdl_amd import core as C
    from oracle import pdlcpu as O

    rng = np.random.default_rng(11)
    res = {"n": n, "check_rows": nrows_check}
    rows = np.sort(rng.choice(n, nrows_check, replace=False))
    for case in ("normal", "lognormal_rows"):
        C.srandom(21)
        a = normal_f32(C, n, n)
        b = normal_f32(C, n, n)
        if case == "lognormal_rows":
            # rows h of A times float(exp(grandom(1,n) * 3)): dims (1,n) broadcast over t
            a = a * (C.grandom(1, n) * 3).exp().convert(C.L.F)
        ah = a.numpy()
        bh = b.numpy()
        truth = ah[rows].astype(np.float64) @ bh.astype(np.float64)
        tn = np.linalg.norm(truth)
        out = {}
        for mode in ("split16", "fp32", "fp8"):
            C.set_matmult_precision(mode)
            c = C.ndarray.new_dense(C.L.F, [n, n])
            C._run("matmult", [a, b, c])          # warm
            C.synchronize()
            e0, e1 = C.Event(), C.Event()
            e0.record()
            for _ in range(reps):
                C._run("matmult", [a, b, c])
            e1.record()
            ms = e0.elapsed_ms(e1) / reps
            got = c.numpy()[rows].astype(np.float64)
            out[mode] = {"ms": round(ms, 3), "TFLOP/s": round(2.0 * n ** 3 / (ms * 1e-3) / 1e12, 1),
                         "rel_err": float(np.linalg.norm(got - truth) / tn)}
        C.set_matmult_precision("split16")
        # PDL's FP32 loop on the sampled rows (sequential float accumulation)
        oa = O.OPdl.from_numpy(np.ascontiguousarray(ah[rows]), O.F)
        ob = O.OPdl.from_numpy(bh, O.F)
        oc = O.OPdl.null()
        t0 = time.perf_counter()
        O.run("matmult", oa, ob, oc)
        dt = time.perf_counter() - t0
        out["pdl_cpu_fp32"] = {"rel_err": float(np.linalg.norm(oc.gather().astype(np.float64) - truth) / tn),
                               "GFLOP/s_1core": round(2.0 * nrows_check * n * n / dt / 1e9, 2)}
        out["split16_vs_fp32_speedup"] = round(out["fp32"]["ms"] / out["split16"]["ms"], 2)
        res[case] = out
    # double on the FP64 path
    m = n // 2
    C.srandom(31)
    a = uniform_pm1(C, m, m, dtype=C.L.D)
    b = uniform_pm1(C, m, m, dtype=C.L.D)
    c = C.ndarray.new_dense(C.L.D, [m, m])
    C._run("matmult", [a, b, c])
    C.synchronize()
    e0, e1 = C.Event(), C.Event()
    e0.record()
    C._run("matmult", [a, b, c])
    e1.record()
    ms = e0.elapsed_ms(e1)
    res["f64"] = {"n": m, "ms": round(ms, 3), "TFLOP/s": round(2.0 * m ** 3 / (ms * 1e-3) / 1e12, 2)}
    return res


def run_small(reps=20):
    """The small CPU-runnable workload: a, b double (2048,2048) uniform[-1,1]
    from PDL's RNG seeded 0, c = a*b + a, s = sumover(c); m = matmult of
    double (512,512) x (512,512) drawn next from the same stream.  Device ms
    per op (HIP events) next to PDL's CPU path (the oracle, autopthread on
    every host core); c must match bit for bit, s and m are scored by their
    relative error against a long-double truth, device and CPU alike.  The
    chain is also run fused (c and s from one pass over a and b)."""
    from pdl_amd import core as C
    from oracle import pdlcpu as O

    N, M = 2048, 512
    D = C.L.D
    C.srandom(0)
    a, b = uniform_pm1(C, N, N, dtype=D), uniform_pm1(C, N, N, dtype=D)
    ma, mb = uniform_pm1(C, M, M, dtype=D), uniform_pm1(C, M, M, dtype=D)
    ah, bh, mah, mbh = a.numpy(), b.numpy(), ma.numpy(), mb.numpy()
    t, c = C.ndarray.new_dense(D, [N, N]), C.ndarray.new_dense(D, [N, N])
    s, m = C.ndarray.new_dense(D, [N]), C.ndarray.new_dense(D, [M, M])
    ops = (("a*b+a", lambda: (C._run("mult", [a, b, t]), C._run("plus", [t, a, c]))),
           ("sumover", lambda: C._run("sumover", [c, s])),
           ("matmult", lambda: C._run("matmult", [ma, mb, m])))
    # the same c and s from one fused pass (pdlhip_chain_prepare): reads a, b once
    fc, fs = C.ndarray.new_dense(D, [N, N]), C.ndarray.new_dense(D, [N])
    ch = C.Chain([a, b], [("mult", 0, 1), ("plus", ("node", 0), 0)], "sumover", value=fc, result=fs)
    ops = ops + (("fused a*b+a -> c, sumover(c) -> s", ch.run),)
    gpu = {}
    for name, f in ops:
        f()
        C.synchronize()
        e0, e1 = C.Event(), C.Event()
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        gpu[name] = round(e0.elapsed_ms(e1) / reps, 4)
    nthr = os.cpu_count() or 1
    oa, ob = O.OPdl.from_numpy(ah, O.D), O.OPdl.from_numpy(bh, O.D)
    oma, omb = O.OPdl.from_numpy(mah, O.D), O.OPdl.from_numpy(mbh, O.D)
    cpu = {}
    t0 = time.perf_counter()
    ot, oc = O.OPdl.null(), O.OPdl.null()
    O.run("mult", oa, ob, ot, nthreads=nthr)
    O.run("plus", ot, oa, oc, nthreads=nthr)
    cpu["a*b+a"] = round((time.perf_counter() - t0) * 1e3, 3)
    t0 = time.perf_counter()
    os_ = O.OPdl.null()
    O.run("sumover", oc, os_, nthreads=nthr)
    cpu["sumover"] = round((time.perf_counter() - t0) * 1e3, 3)
    t0 = time.perf_counter()
    om = O.OPdl.null()
    O.run("matmult", oma, omb, om, nthreads=nthr)
    cpu["matmult"] = round((time.perf_counter() - t0) * 1e3, 3)
    cg = oc.gather()
    s_truth = cg.astype(np.longdouble).sum(axis=-1).astype(np.float64)
    m_truth = O.truth_matmult_ld(mah, mbh)
    rel = lambda x, tr: float(np.linalg.norm(x - tr) / np.linalg.norm(tr))
    return {"workload": "srandom(0); a,b = random(2048,2048)*2-1; c=a*b+a, s=sumover(c); matmult of the next two "
                        "random(512,512)*2-1",
            "gpu_ms": gpu, "cpu_ms": cpu, "cpu_threads": nthr,
            "c_bit_exact": bool(np.array_equal(c.numpy().view(np.uint64), cg.view(np.uint64))),
            "fused_c_bit_exact": bool(np.array_equal(fc.numpy().view(np.uint64), cg.view(np.uint64))),
            "fused_s_rel_err": rel(fs.numpy(), s_truth),
            "s_rel_err": {"gpu": rel(s.numpy(), s_truth), "pdl_cpu": rel(os_.gather(), s_truth)},
            "m_rel_err": {"gpu": rel(m.numpy(), m_truth), "pdl_cpu": rel(om.gather(), m_truth)}}


def run_ragged_matmult(count=4096, lo=8, hi=256, seed=0, reps=5):
    """4096 independent square products, n ~ U{8..256}, float32 and float64
    entries uniform[-1,1]: the list is prepared once (MatmultList: planning,
    item table, autotuned tile), then timed as one persistent work-queue
    launch vs one launch per product; plus a broadcast (64,64)x(64,64)
    matmult over 65536."""
    from pdl_amd import core as C

    C.srandom(seed)
    ns = lo + np.floor(C.random(count).numpy() * (hi - lo + 1)).astype(np.int64)   # U{lo..hi}
    flops = float(np.sum(2.0 * ns.astype(np.float64) ** 3))
    out = {"count": count, "n_range": [lo, hi], "flop": flops,
           "bytes_f32": int(np.sum(3 * ns.astype(np.int64) ** 2) * 4)}
    tot = int(np.sum(ns.astype(np.int64) ** 2))
    for t, name in ((C.L.F, "f32"), (C.L.D, "f64")):
        A = uniform_pm1(C, tot, dtype=t)
        B = uniform_pm1(C, tot, dtype=t)
        Cb = C.ndarray.new_dense(t, [tot])
        As, Bs, Cs, off = [], [], [], 0
        for n in ns:
            n = int(n)
            As.append(C.view_into(A, [n, n], off)); Bs.append(C.view_into(B, [n, n], off))
            Cs.append(C.view_into(Cb, [n, n], off))
            off += n * n
        C.synchronize()
        t0 = time.perf_counter()
        ml = C.MatmultList(As, Bs, Cs)
        C.synchronize()
        res = {"prepare_ms": round((time.perf_counter() - t0) * 1e3, 1)}
        got = {}
        for mode, r in (("queue", reps), ("each", 2)):
            ml.launch(mode)
            C.synchronize()
            e0, e1 = C.Event(), C.Event()
            e0.record()
            for _ in range(r):
                ml.launch(mode)
            e1.record()
            ms = e0.elapsed_ms(e1) / r
            got[mode] = Cb.numpy()
            res[mode] = {"ms": round(ms, 3), "TFLOP/s": round(flops / (ms * 1e-3) / 1e12, 2)}
        res["queue_vs_each"] = round(res["each"]["ms"] / res["queue"]["ms"], 2)
        res["queue_equals_each"] = bool(np.array_equal(got["queue"], got["each"]))
        # the one-call form (planning included), as a PDL loop would call it
        C.matmult_list(As, Bs, Cs, mode="queue")
        C.synchronize()
        t0 = time.perf_counter()
        C.matmult_list(As, Bs, Cs, mode="queue")
        C.synchronize()
        res["unprepared_call_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
        ml.free()
        out[name] = res
    # broadcast matmult (64,64) x (64,64) over a broadcast dim of 65536
    a = uniform_pm1(C, 64, 64, 65536, dtype=C.L.F)
    b = uniform_pm1(C, 64, 64, 65536, dtype=C.L.F)
    c = C.ndarray.new_dense(C.L.F, [64, 64, 65536])
    C._run("matmult", [a, b, c])
    C.synchronize()
    e0, e1 = C.Event(), C.Event()
    e0.record()
    C._run("matmult", [a, b, c])
    e1.record()
    ms = e0.elapsed_ms(e1)
    nbytes = 3 * 64 * 64 * 65536 * 4
    out["broadcast_64x64x65536_f32"] = {"ms": round(ms, 3), "TFLOP/s": round(2 * 64 ** 3 * 65536 / (ms * 1e-3) / 1e12, 2),
                                        "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    return out


RAGGED_CHAIN = [("minus", 0, 1), ("power", ("node", 0), 2), ("mult", ("node", 1), 3), ("mult", ("node", 2), 4)]


def run_ragged_chain(nvec=100000, reps=100, loop_sample=1000, seed=0):
    """100000 float32 vectors, lengths lognormal(8,1.5) clipped to [1,2^20]:
    per vector s = sumover((a-b)**2 * w * m) with a byte mask m, and the
    int64 cumusumover of an int64 vector of the same length.  The fused chain
    and the scan are captured in one HIP graph and replayed `reps` times
    (p50/p99 per replay from HIP events between replays); compared with
    per-op launches over the packed list (4 binops writing temporaries, then
    the ragged sumover and scan) and with PDL's loop of eager ops per vector
    (timed on the first `loop_sample` vectors, scaled to the list)."""
    from pdl_amd import core as C

    F, D, B, LL = C.L.F, C.L.D, C.L.B, C.L.LL
    rng = np.random.default_rng(seed)
    C.srandom(seed)
    # lengths lognormal(mean 8, sigma 1.5) in numpy's sense: exp(8 + 1.5 N(0,1))
    lens = np.clip(np.round(np.exp(8.0 + 1.5 * C.grandom(nvec).numpy())), 1, 2 ** 20).astype(np.int64)
    tot = int(lens.sum())
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
    A = normal_f32(C, tot)
    Bv = normal_f32(C, tot)
    W = C.random(tot, dtype=F)
    M = (C.random(tot) * 2).convert(B)                                  # 0/1 mask
    I = (C.random(tot) * 2 ** 41).convert(LL) - 2 ** 40
    two = C.topdl(2)
    views = lambda big: [C.view_into(big, [int(n)], int(o)) for n, o in zip(lens, offs)]
    va, vb, vw, vm, vi = views(A), views(Bv), views(W), views(M), views(I)
    items = [[a, b, two, w, m] for a, b, w, m in zip(va, vb, vw, vm)]
    t0 = time.perf_counter()
    rs = C.Ragged(items, RAGGED_CHAIN, "sumover")
    rc = C.Ragged([[x] for x in vi], [], "cumusumover")
    prep_s = time.perf_counter() - t0

    def both():
        rs.run()
        rc.run()

    g = C.Graph(both)
    for _ in range(3):
        g.launch()
    C.synchronize()
    ev = [C.Event() for _ in range(reps + 1)]
    ev[0].record()
    for k in range(reps):
        g.launch()
        ev[k + 1].record()
    C.synchronize()
    per = np.array([ev[k].elapsed_ms(ev[k + 1]) for k in range(reps)])
    total_ms = ev[0].elapsed_ms(ev[reps]) / reps
    nbytes = tot * (4 + 4 + 4 + 1) + nvec * 4 + tot * 8 * 2
    out = {"vectors": nvec, "elements": tot, "max_len": int(lens.max()), "bytes_per_replay": nbytes,
           "replays": reps, "prepare_s": round(prep_s, 2),
           "graph": {"ms_mean": round(total_ms, 3), "p50_ms": round(float(np.percentile(per, 50)), 3),
                     "p99_ms": round(float(np.percentile(per, 99)), 3),
                     "GB/s": round(nbytes / (total_ms * 1e-3) / 1e9, 1),
                     "frac": round(nbytes / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}}
    # checks: int64 scans bit-exact against numpy (wrapping adds) on sampled
    # vectors incl. the longest; sums against a float64 sum of float32 chain values
    check_idx = sorted(set([int(np.argmax(lens))] + [int(i) for i in rng.integers(0, nvec, 24)]))
    scan_ok, worst = True, 0.0
    for i in check_idx:
        x = vi[i].numpy()
        scan_ok &= bool(np.array_equal(np.cumsum(x, dtype=np.int64), rc.outputs[i].numpy()))
        z = (va[i].numpy() - vb[i].numpy()) ** 2 * vw[i].numpy() * vm[i].numpy().astype(np.float32)
        truth = float(np.sum(z.astype(np.float64)))
        got = float(rs.outputs[i].numpy())
        worst = max(worst, abs(got - truth) / max(abs(truth), 1e-300))
    out["checks"] = {"i64_cumusumover_exact": scan_ok, "sum_max_rel_err": worst, "sampled": len(check_idx)}
    # per-op launches over the packed list
    z1, z2, z3, z4 = (C.ndarray.new_dense(F, [tot]) for _ in range(4))
    ru = C.Ragged([[v] for v in views(z4)], [], "sumover")

    def unfused():
        C._run("minus", [A, Bv, z1])
        C._run("power", [z1, two, z2])
        C._run("mult", [z2, W, z3])
        C._run("mult", [z3, M, z4])
        ru.run()
        rc.run()

    unfused()
    C.synchronize()
    ev2 = [C.Event() for _ in range(6)]
    ev2[0].record()
    for k in range(5):
        unfused()
        ev2[k + 1].record()
    C.synchronize()
    uper = np.array([ev2[k].elapsed_ms(ev2[k + 1]) for k in range(5)])
    out["per_op_launches"] = {"ms_mean": round(float(uper.mean()), 3), "p50_ms": round(float(np.median(uper)), 3),
                              "graph_speedup": round(float(uper.mean()) / total_ms, 2)}
    same = all(float(ru.outputs[i].numpy()) == float(rs.outputs[i].numpy()) or
               abs(float(ru.outputs[i].numpy()) - float(rs.outputs[i].numpy())) <=
               4 * 2.0 ** -24 * abs(float(rs.outputs[i].numpy())) for i in check_idx)
    out["checks"]["fused_vs_per_op_within_4ulp"] = bool(same)
    del z1, z2, z3, z4
    # PDL's loop: eager ops per vector (temporaries allocated per op)
    ns = min(loop_sample, nvec)
    C.synchronize()
    t0 = time.perf_counter()
    for i in range(ns):
        z = (va[i] - vb[i]) ** 2 * vw[i] * vm[i]
        s = z.sumover()
        c = vi[i].cumusumover()
    C.synchronize()
    loop_s = (time.perf_counter() - t0) * nvec / ns
    out["per_vector_loop"] = {"sample_vectors": ns, "s_scaled_to_list": round(loop_s, 3),
                              "graph_speedup": round(loop_s * 1e3 / total_ms, 1)}
    g.destroy()
    rs.free(); rc.free(); ru.free()
    return out


def run_single(args):
    """one op alone, for counter runs of its kernel"""
    from pdl_amd import core as C

    C.init(0)
    if args.only == "ragged":
        print(json.dumps(run_ragged_chain()))
        return
    if args.only == "small":
        print(json.dumps(run_small()))
        return
    if args.only == "matmult":
        print(json.dumps({"matmult": run_matmult(), "ragged_matmult": run_ragged_matmult()}))
        return
    if args.only == "lists":
        print(json.dumps({"ragged_matmult": run_ragged_matmult()}))
        return
    if args.only == "listq":
        # the two prepared lists' queue launches alone, for counter runs
        C.srandom(0)
        ns = 8 + np.floor(C.random(4096).numpy() * 249).astype(np.int64)
        tot = int(np.sum(ns ** 2))
        for t in (C.L.F, C.L.D):
            A, B = uniform_pm1(C, tot, dtype=t), uniform_pm1(C, tot, dtype=t)
            Cb = C.ndarray.new_dense(t, [tot])
            As, Bs, Cs, off = [], [], [], 0
            for n in ns:
                n = int(n)
                As.append(C.view_into(A, [n, n], off)); Bs.append(C.view_into(B, [n, n], off))
                Cs.append(C.view_into(Cb, [n, n], off))
                off += n * n
            ml = C.MatmultList(As, Bs, Cs)
            for _ in range(args.steps):
                ml.launch("queue")
            C.synchronize()
            ml.free()
        print(json.dumps({"only": args.only, "steps": args.steps}))
        return
    if args.only in ("mm16", "mm8"):
        C.set_matmult_precision("split16" if args.only == "mm16" else "fp8")
        C.srandom(21)
        a = normal_f32(C, 8192, 8192)
        bm = normal_f32(C, 8192, 8192)
        cm = C.ndarray.new_dense(C.L.F, [8192, 8192])
        for _ in range(args.steps):
            C._run("matmult", [a, bm, cm])
        C.synchronize()
        print(json.dumps({"only": args.only, "steps": args.steps}))
        return
    if args.only == "copy":
        C.srandom(1)
        src = C.random(256, 1024, 1024, dtype=C.L.F)                      # 1 GiB
        dst = C.ndarray.new_dense(C.L.F, src.dims)
        f = lambda: C._run("assgn", [src, dst])
    else:
        C.srandom(1)
        x = normal_f32(C, 4096, 4096, 16)
        y = normal_f32(C, 4096, 16)
        z = C.ndarray.new_dense(C.L.F, [4096, 4096, 16])
        if args.only == "binop":
            xv, yv = x.xchg(0, 1), y.dummy(1)
            f = lambda: C._run("mult", [xv, yv, z])
        else:
            o = C.ndarray.new_dense(C.L.F, [4096, 16])
            f = lambda: C._run(args.only, [x, o])
    for _ in range(args.steps):
        f()
    C.synchronize()
    print(json.dumps({"only": args.only, "steps": args.steps}))


def host_cores():
    """the host threads this process may use (its affinity set, at most 16:
    the CPU share of one GPU on the measurement hosts) and the CPU model"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return min(n, 16), model


def run_cpu_baseline(frac_k=2, min_seconds=10.0, min_runs=5):
    """The oracle (the C restatement of PDL's CPU path) on a bounded sample of
    the same workload: x (4096,4096,frac_k) instead of 16 k-planes, the int
    arrays at 1/4 of their rows; PDL's autopthread rule over the host cores.
    One warm-up pass, then passes until `min_seconds` and `min_runs` are
    reached; the rate is taken from the median pass."""
    from oracle import pdlcpu as O

    nthr, model = host_cores()
    NX, NY, NK = 4096, 4096, frac_k
    # the same expressions as the device inputs, from the oracle's stream
    u = O.random_from(NX * NY * NK + NY * NK, 1)
    g = O.ndtri(u).astype(np.float32)
    x = g[:NX * NY * NK].reshape(NK, NY, NX)
    y = g[NX * NY * NK:].reshape(NK, NY)
    r = O.random_from(4096 * 1024 + 2 * 8192 * 2048, 2)
    i32h = ((r[:4096 * 1024] * 2097153).astype(np.int32) - 1048576).reshape(1024, 4096)
    v = ((r[4096 * 1024:4096 * 1024 + 8192 * 2048] * 65535).astype(np.int32) - 32767).astype(np.int16)
    v[r[4096 * 1024 + 8192 * 2048:] < 0.01] = -32768
    i16h = v.reshape(2048, 8192)
    i16h[17] = -32768
    px = O.OPdl.from_numpy(x, O.F).xchg(0, 1)
    py = O.OPdl.from_numpy(y, O.F).dummy(1)
    p32 = O.OPdl.from_numpy(i32h, O.L)
    p16 = O.OPdl.from_numpy(i16h, O.S, badflag=True)
    nb = algorithmic_bytes(NX, NY, NK, 4096, 1024, 8192, 2048)

    def one_pass():
        t0 = time.perf_counter()
        z = O.OPdl.null()
        O.run("mult", px, py, z, nthreads=nthr)
        for op in ("sumover", "average", "minimum", "maximum"):
            O.run(op, z, O.OPdl.null(), nthreads=nthr)
        O.run("sumover", p32, O.OPdl.null(), nthreads=nthr)
        O.run("cumusumover", p32, O.OPdl.null(), nthreads=nthr)
        O.run("sumover", p16, O.OPdl.null(), nthreads=nthr)
        return time.perf_counter() - t0

    one_pass()
    times = []
    while len(times) < min_runs or sum(times) < min_seconds:
        times.append(one_pass())
    med = float(np.median(times))
    return {"value": round(sum(nb.values()) / med / 1e9, 2), "unit": "GB/s", "cores": nthr, "kind": "port",
            "sample": "the same chain at x (4096,4096,%d), int32 (4096,1024), int16 (8192,2048): median of %d passes "
                      "(%.1f s in all) on %s" % (NK, len(times), sum(times), model)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-matmult", action="store_true")
    ap.add_argument("--no-ragged", action="store_true")
    ap.add_argument("--only", default=None, help="run one op alone (binop, sumover, copy ...) for counters")
    args = ap.parse_args()
    if args.only:
        run_single(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    r = run_gpu(args)
    per = r["per_op"][r["dominant"]]
    out = {
        "metric": METRIC, "value": round(r["value"], 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32+i16", "data": "synthetic",
        "config": {"workload": "broadcast+reductions at HBM scale: z=x->xchg(0,1)*y->dummy(1) f32 (4096,4096,16); "
                               "sumover/average/minimum/maximum dim 0; i32 (4096,4096) sumover+cumusumover; "
                               "i16 (8192,8192) 1% BAD sumover",
                   "bytes_per_step": r["total_bytes"], "frac_of_hbm_peak": round(r["value"] / HBM_PEAK_GBS, 3)},
        "roofline": {"bound": "hbm", "achieved": per["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": per["frac"], "traffic": measured_traffic(r["dominant"])[0],
                     "traffic_kernel": measured_traffic(r["dominant"])[1], "kernel": r["dominant"]},
        "per_op": r["per_op"], "checks": r["check"],
    }
    out["small"] = run_small()
    if not args.no_matmult:
        out["matmult"] = run_matmult()
        out["ragged_matmult"] = run_ragged_matmult()
    if not args.no_ragged:
        out["ragged_chain"] = run_ragged_chain()
    if not args.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
