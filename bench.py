#!/usr/bin/env python3
"""Benchmark of the compressed chunk-blocked matmul path on one MI355X.

Default workload (the configuration the headline metric is quoted on): plain
FP32 GEMV y = A @ x, A 20000x20000 float32 with A[i,j] = sin(0.001 i) +
cos(0.0007 j) + N(0, 1e-3), x 20000 float32 U[-1,1), A chunks 2000x2000 /
blocks 250x250, x and y chunks 2000 / blocks 250, LZ4 clevel 5 + bitshuffle.
One step = one iarray_linalg_matmul call on operands already resident in HBM
(decode of every A block + GEMV + C re-shuffle + host LZ4 write of y).

Prints ONE JSON line.  The roofline object describes the dominant kernel of
the timed steps, timed with HIP events on the library stream; cpu_baseline is
the oracle's restatement of the reference CPU path (Blosc2-style block
threads, liblz4 decode, MKL single-threaded per block) on this host.
"""
import argparse
import json
import os
import struct
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "iron-array_amd"))

METRIC = "matmul GFLOP/s incl. block decode, %roofline; mixed vs plain prec at fixed err"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0, "i8": 5000.0}


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="gemv20k",
                   choices=["gemv20k", "gemm2k_f64", "gemm8k_f32", "gemm6k_f64", "batch512", "hbm_copy"])
    p.add_argument("--slices", type=int, default=0, help="Ozaki slice count (0 = autotuned to 1e-13)")
    p.add_argument("--precision", default="plain", choices=["plain", "split_bf16", "ozaki_int8"])
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=0)
    return p.parse_args()


def stream_bytes(chunks):
    """(compressed bytes of LZ4 streams, decoded bytes of LZ4 streams, raw bytes)
    of a list of our own Blosc2 chunks (version-5 layout)."""
    lz_in = lz_out = raw = 0
    for c in chunks:
        flags, ts = c[2], c[3]
        nbytes, bsize = struct.unpack("<ii", c[4:12])
        if (c[31] >> 4) & 7 or flags & 0x2:
            continue
        nblocks = nbytes // bsize
        nstreams = 1 if flags & 0x10 else ts
        neb = bsize // nstreams
        for b in range(nblocks):
            p = struct.unpack("<i", c[32 + 4 * b:36 + 4 * b])[0]
            for _ in range(nstreams):
                cs = struct.unpack("<i", c[p:p + 4])[0]
                p += 4
                if cs == neb:
                    raw += cs
                elif cs > 0:
                    lz_in += cs
                    lz_out += neb
                p += max(cs, 0)
    return lz_in, lz_out, raw


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the library runs one product on one "
              "GPU and has no path over several", file=sys.stderr)
        sys.exit(2)
    if args.workload == "hbm_copy":
        hbm_copy(args)
        return
    import pyiarray as ia

    name, cus, arch = ia.device_info()
    print(f"# device {name} {arch} {cus} CUs", file=sys.stderr, flush=True)
    rng = np.random.default_rng(1001)
    t_setup = time.time()
    if args.workload == "gemv20k":
        M = K = 20000
        ctx = ia.Context(clevel=5, filters=ia.COMP_BITSHUFFLE)
        A = rng.standard_normal((M, K), dtype=np.float32)
        A *= np.float32(1e-3)
        A += np.sin(0.001 * np.arange(M, dtype=np.float64)).astype(np.float32)[:, None]
        A += np.cos(0.0007 * np.arange(K, dtype=np.float64)).astype(np.float32)[None, :]
        x = rng.uniform(-1, 1, K).astype(np.float32)
        cA = ia.Container.from_numpy(ctx, A, (2000, 2000), (250, 250))
        cx = ia.Container.from_numpy(ctx, x, (2000,), (250,))
        del A
        out_chunks, out_blocks = (2000,), (250,)
        flops = 2.0 * M * K
        dtype = "f32"
        config = {"workload": "gemv 20000x20000 f32, A chunks 2000x2000 blocks 250x250, LZ4 clevel5 + bitshuffle",
                  "M": M, "K": K, "chunkshape": [2000, 2000], "blockshape": [250, 250], "codec": "lz4",
                  "clevel": 5, "filter": "bitshuffle"}
        step = lambda: ia.matmul(cA, cx, out_chunks, out_blocks)  # noqa: E731
    elif args.workload == "batch512":
        ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
        dims = rng.integers(64, 1537, (512, 3))
        pairs, cs_l, bs_l = [], [], []
        flops = 0.0
        for M_, N_, K_ in dims:
            M_, N_, K_ = int(M_), int(N_), int(K_)
            a = rng.uniform(-1, 1, (M_, K_)).astype(np.float32)
            b = rng.uniform(-1, 1, (K_, N_)).astype(np.float32)
            ca_ = ia.Container.from_numpy(ctx, a, (min(M_, 512), min(K_, 512)), (min(M_, 512, 128), min(K_, 128)))
            cb_ = ia.Container.from_numpy(ctx, b, (min(K_, 512), min(N_, 512)), (min(K_, 128), min(N_, 128)))
            pairs.append((ca_, cb_))
            cs_l.append((min(M_, 512), min(N_, 512)))
            bs_l.append((min(M_, 128), min(N_, 128)))
            flops += 2.0 * M_ * N_ * K_
        cA, cx = pairs[0]
        step = lambda: ia.matmul_batch(pairs, cs_l, bs_l)  # noqa: E731
        M = K = 0
        dtype = "f32"
        config = {"workload": "512 independent f32 matmuls, M,N,K ~ U{64..1536} (seed 1001), chunks min(dim,512), "
                              "blocks 128x128 with ragged edges, LZ4 clevel5 + shuffle, one graph replay per phase of products",
                  "products": 512, "total_flop": flops}
        batch_spec = (pairs, cs_l, bs_l)
    else:
        spec = {"gemm2k_f64": (2000, np.float64, (500, 500), (125, 125), 5),
                "gemm8k_f32": (8192, np.float32, (1024, 1024), (256, 256), 3),
                "gemm6k_f64": (6000, np.float64, (1500, 1500), (250, 250), 5)}[args.workload]
        n, dt, cs, bs, clevel = spec
        ctx = ia.Context(clevel=clevel, filters=ia.COMP_SHUFFLE)
        if args.workload == "gemm2k_f64":
            a = rng.uniform(-1, 1, (n, n))
            b = rng.uniform(-1, 1, (n, n))
        elif args.workload == "gemm8k_f32":
            a = rng.standard_normal((n, n), dtype=np.float32)
            b = rng.standard_normal((n, n), dtype=np.float32)
        else:
            a = np.where(rng.random((n, n)) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, (n, n))
            b = np.where(rng.random((n, n)) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, (n, n))
        a = a.astype(dt)
        b = b.astype(dt)
        cA = ia.Container.from_numpy(ctx, a, cs, bs)
        cx = ia.Container.from_numpy(ctx, b, cs, bs)
        prec = {"plain": ia.PLAIN, "split_bf16": ia.SPLIT_BF16, "ozaki_int8": ia.OZAKI_INT8}[args.precision]
        step = lambda: ia.matmul(cA, cx, cs, bs, precision=prec, slices=args.slices)  # noqa: E731
        err_sample = (a, b)
        M = K = n
        flops = 2.0 * n ** 3
        dtype = "f64" if dt == np.float64 else "f32"
        config = {"workload": f"gemm {n}x{n}x{n} {dtype}, chunks {cs[0]}x{cs[1]} blocks {bs[0]}x{bs[1]}, "
                              f"LZ4 clevel{clevel} + shuffle, {args.precision}",
                  "M": n, "N": n, "K": n, "precision": args.precision}
    print(f"# setup (generate + host encode) {time.time() - t_setup:.1f}s", file=sys.stderr, flush=True)
    nb, cb = cA.info()

    sync = ia.device_sync

    ia.prof_enable(True)
    for _ in range(args.warmup):
        step()
    sync()
    ia.prof_reset()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c = step()
    sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    prof = ia.prof_entries()
    enc_s, _ = ia.last_matmul_times()
    if args.precision == "ozaki_int8":
        S, _ = ia.last_ozaki_slices()
        config["ozaki_slices"] = S
        config["ozaki_slices_chosen_by"] = "argument" if args.slices else "sampled-error autotune (target 1e-13)"
    ia.prof_enable(False)
    print(f"# kernels: {json.dumps({k: [v[0], round(v[1] / max(v[0], 1), 4)] for k, v in prof.items()})}",
          file=sys.stderr, flush=True)

    # roofline of the dominant kernel.  The library's launch timings cover its
    # last call, i.e. one step (row-panel GEMMs launch once a panel, Ozaki
    # once a level and panel).  The batch runs as graph replays, so its
    # kernels are timed on one more, eager replay of the same launches.
    if args.workload == "batch512":
        os.environ["IARRAY_BATCH_EAGER"] = "1"
        ia.prof_enable(True)
        ia.prof_reset()
        step()
        sync()
        prof = {k: v for k, v in ia.prof_entries().items()}
        ia.prof_enable(False)
        del os.environ["IARRAY_BATCH_EAGER"]
        print(f"# kernels (eager replay): {json.dumps({k: [v[0], round(v[1] / max(v[0], 1), 4)] for k, v in prof.items()})}",
              file=sys.stderr, flush=True)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dname, (dn, dms) = dom
    avg_s = dms / dn / 1e3
    per_step = dn  # launches of this kernel in one step
    isz = 4 if dtype == "f32" else 8
    if dname.startswith("lz4_streams"):
        ops = [o for pr in pairs for o in pr] if args.workload == "batch512" else (
            [cA] if args.workload == "gemv20k" else [cA, cx])
        alg = 0
        for o in ops:
            li, lo, _ = stream_bytes(o.chunks())
            alg += li + lo
        alg /= per_step
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
    elif dname.startswith("lz4_encode"):
        # the GPU share of the C write: filtered stream bytes read + finished chunk bytes written
        gin, gout = ia.last_gpu_encode_bytes()
        alg = (gin + gout) / per_step
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
    elif dname.startswith("gemv_fused"):
        lz_in, lz_out, raw = stream_bytes(cA.chunks())
        alg = (lz_out + raw + K * isz + (M // 250) * M * isz) / per_step
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
    elif dname.startswith("gemm"):
        # matrix-core work of a step: 3 bf16 products for split-bf16, one int8
        # product per slice pair of the kept levels (S(S+1)/2) for Ozaki
        npair = 1
        if args.precision == "split_bf16":
            npair = 3
        elif args.precision == "ozaki_int8":
            S = config["ozaki_slices"]
            npair = S * (S + 1) // 2
        alg = flops * npair / per_step
        key = {"plain": dtype, "split_bf16": "bf16", "ozaki_int8": "i8"}[args.precision]
        bound, unit, peak, achieved = "mfma", "TFLOP/s", PEAK_TFLOPS[key], alg / avg_s / 1e12
    else:
        alg = nb / per_step
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
    roof = {"bound": bound, "kernel": dname, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": measured_traffic(args, dname),
            "algorithmic_bytes_or_flops_per_launch": alg, "avg_launch_ms": round(avg_s * 1e3, 4),
            "launches_per_step": per_step}

    errors = None
    if args.workload.startswith("gemm"):
        errors = error_metrics(c, *err_sample)
    cpu = None
    if not args.no_cpu:
        if args.workload == "batch512":
            cpu = cpu_baseline_batch(args, batch_spec)
        else:
            cpu = cpu_baseline(args, ia, cA, cx, config, flops)
    out = {"metric": METRIC, "value": round(flops / (ms / 1e3) / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded numpy generator)",
           "config": config, "roofline": roof, "cpu_baseline": cpu, "errors_vs_fp64_oracle": errors,
           "compressed_bytes_A": cb, "decoded_bytes_A": nb,
           "host_encode_ms_last_step": round(enc_s * 1e3, 3),
           "compressed_GBps": round(cb / (ms / 1e3) / 1e9, 2), "decoded_GBps": round(nb / (ms / 1e3) / 1e9, 2)}
    print(json.dumps(out), flush=True)


def hbm_copy(args):
    """Calibration workload: a plain 1 GiB device-to-device copy (torch's
    elementwise copy kernel), the reference for scaling the FETCH_SIZE and
    WRITE_SIZE counters to bytes."""
    import torch

    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    print(json.dumps({"metric": "hbm copy (read + write) GB/s", "value": round(2 * 4 * n / ms / 1e6, 1),
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(ms, 4), "bytes_read": 4 * n, "bytes_written": 4 * n}), flush=True)


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` on this workload, from the FETCH_SIZE
    and WRITE_SIZE counter runs recorded in profiles/hbm_traffic.json (each
    already scaled by the copy-kernel calibration); None when no record
    matches the workload, precision and kernel."""
    path = os.path.join(REPO, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    rec = table.get(f"{args.workload}/{args.precision}/{kernel}")
    return None if rec is None else rec["bytes_per_launch"]


def error_metrics(c, a, b, nsample=4096):
    """Errors of C against a double-double oracle on a seeded sample of entries:
    elementwise max |C-C*|/|C*|, componentwise max |C-C*|/(|A||B|), and the
    sample's normwise ||C-C*||/||C*||."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    rng = np.random.default_rng(7)
    M, N = a.shape[0], b.shape[1]
    rows = rng.integers(0, M, nsample)
    cols = rng.integers(0, N, nsample)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    ref = orc.dd_gemm_sample(a64, b64, rows, cols)
    full = c.to_numpy().astype(np.float64)
    got = full[rows, cols]
    den = np.einsum("ij,ji->i", np.abs(a64[rows]), np.abs(b64[:, cols]))
    err = np.abs(got - ref)
    return {"samples": nsample, "elementwise_max": float(np.max(err / np.maximum(np.abs(ref), 1e-300))),
            "componentwise_max": float(np.max(err / den)),
            "normwise": float(np.linalg.norm(err) / np.linalg.norm(ref))}


def cpu_baseline(args, ia, cA, cB, config, flops):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    blas = orc.blas_name() if orc.set_blas(1) else "plain-c"
    orc.set_lz4(1)
    a = orc.Array(cA.chunks(), cA.shape, cA_chunkshape(config, cA), cA_blockshape(config, cA), cA.dtype.itemsize)
    if config["workload"].startswith("gemv"):
        b = orc.Array(cB.chunks(), cB.shape, (2000,), (250,), cB.dtype.itemsize)
        fn = lambda: orc.gemv_blosc(a, b, (2000,), (250,), nthreads=threads)  # noqa: E731
        filt = 2
        cc, cbk = (2000,), (250,)
    else:
        b = orc.Array(cB.chunks(), cB.shape, cA_chunkshape(config, cA), cA_blockshape(config, cA), cB.dtype.itemsize)
        cc, cbk = cA_chunkshape(config, cA), cA_blockshape(config, cA)
        fn = lambda: orc.gemm_blosc(a, b, cc, cbk, nthreads=threads)  # noqa: E731
        filt = 1
    t0 = time.perf_counter()
    c = fn()
    t_mm = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.encode_dense(np.ascontiguousarray(c), cc, cbk, 5, filt, threads)
    t_enc = time.perf_counter() - t0
    total = t_mm + t_enc
    return {"value": round(flops / total / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"the full workload once (matmul {t_mm:.2f}s + C encode {t_enc:.2f}s), blas={blas}, "
                      f"lz4=liblz4"}


def cpu_baseline_batch(args, spec, nsample=48):
    """The reference runs the 512 products as sequential iarray_linalg_matmul
    calls; time the first `nsample` of them (decode + blocked GEMM + C encode,
    block-parallel over the host cores) and report their GFLOP/s."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    pairs, cs_l, bs_l = spec
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    blas = orc.blas_name() if orc.set_blas(1) else "plain-c"
    orc.set_lz4(1)
    flops = 0.0
    t0 = time.perf_counter()
    for (ca, cb), cs, bs in list(zip(pairs, cs_l, bs_l))[:nsample]:
        M_, K_ = ca.shape
        N_ = cb.shape[1]
        a = orc.Array(ca.chunks(), ca.shape, (min(M_, 512), min(K_, 512)), (min(M_, 512, 128), min(K_, 128)), 4)
        b = orc.Array(cb.chunks(), cb.shape, (min(K_, 512), min(N_, 512)), (min(K_, 128), min(N_, 128)), 4)
        c = orc.gemm_blosc(a, b, cs, bs, nthreads=threads)
        orc.encode_dense(np.ascontiguousarray(c), cs, bs, 5, 1, threads)
        flops += 2.0 * M_ * N_ * K_
    total = time.perf_counter() - t0
    return {"value": round(flops / total / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"the first {nsample} of the 512 products run sequentially ({total:.2f}s, "
                      f"{flops:.3g} flop), blas={blas}, lz4=liblz4"}


def cA_chunkshape(config, c):
    return tuple(config.get("chunkshape", [])) or {8192: (1024, 1024), 6000: (1500, 1500), 2000: (500, 500)}[c.shape[0]]


def cA_blockshape(config, c):
    return tuple(config.get("blockshape", [])) or {8192: (256, 256), 6000: (250, 250), 2000: (125, 125)}[c.shape[0]]


if __name__ == "__main__":
    main()
