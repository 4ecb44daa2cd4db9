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
threads, liblz4 decode, MKL single-threaded per block) on this host.  The
default run also carries `mixed_vs_plain`: the 8192^3 float32 product in its
plain and split-BF16 forms and the 6000^3 float64 product in its plain and
Ozaki-INT8 forms, each timed and checked against a double-double oracle.

The workload generators (`make_workload`) are shared with the GPU tests, so
the tests check exactly the operands measured here.
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
WORKLOADS = ["gemv20k", "gemm2k_f64", "gemm8k_f32", "gemm6k_f64", "batch512"]
PRECISIONS = {"plain": 0, "split_bf16": 1, "ozaki_int8": 2, "ozaki_crt": 3}
# matrix-core products of 2MNK each per step, by form (count = Ozaki slices or moduli)
PEAK_KEY = {"plain": None, "split_bf16": "bf16", "ozaki_int8": "i8", "ozaki_crt": "i8"}


def npairs(precision, count):
    return {"plain": 1, "split_bf16": 3, "ozaki_int8": (count * (count + 1) // 2) if count else 1,
            "ozaki_crt": count or 1}[precision]


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="gemv20k", choices=WORKLOADS + ["hbm_copy"])
    p.add_argument("--slices", type=int, default=0, help="Ozaki slice count (0 = autotuned to 1e-13)")
    p.add_argument("--precision", default="plain", choices=list(PRECISIONS))
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-mixed", action="store_true", help="skip the mixed-vs-plain block")
    p.add_argument("--mixed-steps", type=int, default=5)
    p.add_argument("--mixed-warmup", type=int, default=5,
                   help="untimed calls per form (the C write's GPU share settles in these)")
    p.add_argument("--cpu-threads", type=int, default=0)
    return p.parse_args()


# ---- workloads -----------------------------------------------------------------------------
def gemv20k_matrix(rng, M=20000, K=20000):
    """A[i,j] = sin(0.001 i) + cos(0.0007 j) + N(0, 1e-3), float32."""
    A = rng.standard_normal((M, K), dtype=np.float32)
    A *= np.float32(1e-3)
    A += np.sin(0.001 * np.arange(M, dtype=np.float64)).astype(np.float32)[:, None]
    A += np.cos(0.0007 * np.arange(K, dtype=np.float64)).astype(np.float32)[None, :]
    return A


def wide_range(rng, shape):
    """sign * 10^u, u ~ U[-3, 3], random sign (float64)."""
    return np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-3, 3, shape)


def make_workload(ia, name, keep_dense=False):
    """Operands of one named workload, compressed into containers.

    Returns a dict: ctx, a, b (containers), a_dense/b_dense (numpy, when
    keep_dense or needed for errors), c_chunks, c_blocks, flops, dtype and
    config; for batch512 `pairs`, `c_chunks_list`, `c_blocks_list` and the
    dense operands of every product."""
    rng = np.random.default_rng(1001)
    if name == "gemv20k":
        M = K = 20000
        ctx = ia.Context(clevel=5, filters=ia.COMP_BITSHUFFLE)
        A = gemv20k_matrix(rng, M, K)
        x = rng.uniform(-1, 1, K).astype(np.float32)
        cA = ia.Container.from_numpy(ctx, A, (2000, 2000), (250, 250))
        cx = ia.Container.from_numpy(ctx, x, (2000,), (250,))
        config = {"workload": "gemv 20000x20000 f32, A chunks 2000x2000 blocks 250x250, LZ4 clevel5 + bitshuffle",
                  "M": M, "K": K, "chunkshape": [2000, 2000], "blockshape": [250, 250], "codec": "lz4",
                  "clevel": 5, "filter": "bitshuffle"}
        return {"ctx": ctx, "a": cA, "b": cx, "a_dense": A, "b_dense": x, "c_chunks": (2000,), "c_blocks": (250,),
                "a_tiling": ((2000, 2000), (250, 250)), "b_tiling": ((2000,), (250,)),
                "flops": 2.0 * M * K, "dtype": "f32", "config": config}
    if name == "batch512":
        ctx = ia.Context(clevel=5, filters=ia.COMP_SHUFFLE)
        dims = rng.integers(64, 1537, (512, 3))
        pairs, cs_l, bs_l, dense = [], [], [], []
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
            dense.append((a, b) if keep_dense else None)
            flops += 2.0 * M_ * N_ * K_
        config = {"workload": "512 independent f32 matmuls, M,N,K ~ U{64..1536} (seed 1001), chunks min(dim,512), "
                              "blocks 128x128 with ragged edges, LZ4 clevel5 + shuffle, one graph replay per phase of products",
                  "products": 512, "total_flop": flops}
        return {"ctx": ctx, "pairs": pairs, "c_chunks_list": cs_l, "c_blocks_list": bs_l, "dense": dense,
                "a": pairs[0][0], "b": pairs[0][1], "flops": flops, "dtype": "f32", "config": config}
    n, dt, cs, bs, clevel = {"gemm2k_f64": (2000, np.float64, (500, 500), (125, 125), 5),
                             "gemm8k_f32": (8192, np.float32, (1024, 1024), (256, 256), 3),
                             "gemm6k_f64": (6000, np.float64, (1500, 1500), (250, 250), 5)}[name]
    ctx = ia.Context(clevel=clevel, filters=ia.COMP_SHUFFLE)
    if name == "gemm2k_f64":
        a = rng.uniform(-1, 1, (n, n))
        b = rng.uniform(-1, 1, (n, n))
    elif name == "gemm8k_f32":
        a = rng.standard_normal((n, n), dtype=np.float32)
        b = rng.standard_normal((n, n), dtype=np.float32)
    else:
        a = wide_range(rng, (n, n))
        b = wide_range(rng, (n, n))
    a, b = a.astype(dt), b.astype(dt)
    cA = ia.Container.from_numpy(ctx, a, cs, bs)
    cB = ia.Container.from_numpy(ctx, b, cs, bs)
    dtype = "f64" if dt == np.float64 else "f32"
    config = {"workload": f"gemm {n}x{n}x{n} {dtype}, chunks {cs[0]}x{cs[1]} blocks {bs[0]}x{bs[1]}, "
                          f"LZ4 clevel{clevel} + shuffle",
              "M": n, "N": n, "K": n}
    return {"ctx": ctx, "a": cA, "b": cB, "a_dense": a, "b_dense": b, "c_chunks": cs, "c_blocks": bs,
            "a_tiling": (cs, bs), "b_tiling": (cs, bs), "flops": 2.0 * n ** 3, "dtype": dtype, "config": config}


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


# ---- accuracy ------------------------------------------------------------------------------
def error_metrics(c_dense, a, b, nsample=4096, seed=7):
    """Errors of C against a double-double oracle (exact dot product rounded
    once) on a seeded sample of entries: elementwise max |C-C*|/|C*|,
    componentwise max |C-C*|/(|A||B|), and the sample's normwise
    ||C-C*||/||C*||.  A 1-D b is a GEMV (rows sampled)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    rng = np.random.default_rng(seed)
    M = a.shape[0]
    if b.ndim == 1:
        rows = np.unique(rng.integers(0, M, nsample))
        a64 = a[rows].astype(np.float64)
        b64 = b.astype(np.float64)[:, None]
        ref = orc.dd_gemm_sample(a64, b64, np.arange(len(rows)), np.zeros(len(rows), np.int64))
        got = c_dense[rows].astype(np.float64)
        den = np.abs(a64) @ np.abs(b64[:, 0])
    else:
        N = b.shape[1]
        rows = rng.integers(0, M, nsample)
        cols = rng.integers(0, N, nsample)
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        ref = orc.dd_gemm_sample(a64, b64, rows, cols)
        got = c_dense[rows, cols].astype(np.float64)
        den = np.einsum("ij,ji->i", np.abs(a64[rows]), np.abs(b64[:, cols]))
    err = np.abs(got - ref)
    return {"samples": int(len(ref)), "elementwise_max": float(np.max(err / np.maximum(np.abs(ref), 1e-300))),
            "componentwise_max": float(np.max(err / np.where(den > 0, den, 1.0))),
            "normwise": float(np.linalg.norm(err) / np.linalg.norm(ref))}


# ---- timing --------------------------------------------------------------------------------
def timed_steps(ia, step, warmup, steps):
    """W untimed steps, then K timed steps bracketed by device syncs; returns
    (ms per step, last result, per-kernel HIP-event timings of the last step)."""
    ia.prof_enable(True)
    for _ in range(warmup):
        step()
    ia.device_sync()
    ia.prof_reset()
    t0 = time.perf_counter()
    c = None
    for _ in range(steps):
        c = step()
    ia.device_sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    prof = ia.prof_entries()
    ia.prof_enable(False)
    return ms, c, prof


def roofline(ia, w, prof, precision, npair=1):
    """Roofline of the dominant kernel (most device time in one step).  For
    a decode-dominated GEMV the algorithmic bytes are those of the step:
    compressed bytes of A + 4 (K + M) (x read, y written); the decoder's own
    scratch traffic is not algorithmic."""
    dname, (dn, dms) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_s = dms / dn / 1e3
    dtype = w["dtype"]
    isz = 4 if dtype == "f32" else 8
    flops = w["flops"]
    note = None
    if dname.startswith("gemm"):
        alg = flops * npair / dn
        key = PEAK_KEY[precision] or dtype
        bound, unit, peak, achieved = "mfma", "TFLOP/s", PEAK_TFLOPS[key], alg / avg_s / 1e12
        note = f"{npair} matrix-core product(s) of 2MNK each per step"
    elif "pairs" in w:
        gin, gout = ia.last_gpu_encode_bytes()
        alg = (gin + gout) / dn
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
        note = "filtered stream bytes read + finished chunk bytes written by the GPU share of the C write"
    elif len(w["b"].shape) == 1:
        M, K = w["config"]["M"], w["config"]["K"]
        _, cb = w["a"].info()
        alg = (cb + isz * (K + M)) / dn
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
        note = "algorithmic bytes = compressed bytes of A + itemsize*(K+M); decoded scratch not counted"
    else:
        ops = [w["a"], w["b"]]
        alg = 0
        for o in ops:
            li, lo, _ = stream_bytes(o.chunks())
            alg += li + lo
        alg /= dn
        bound, unit, peak, achieved = "hbm", "GB/s", HBM_PEAK_GBS, alg / avg_s / 1e9
        note = "compressed LZ4 bytes read + decoded bytes written"
    return {"bound": bound, "kernel": dname, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "algorithmic_bytes_or_flops_per_launch": alg,
            "avg_launch_ms": round(avg_s * 1e3, 4), "launches_per_step": dn, "note": note}


def profile_traffic(workload, precision, kernel):
    """HBM bytes per launch of `kernel` on this workload from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (rocprofv3 --pmc, scaled by the
    copy-kernel calibration) in profiles/hbm_traffic.json; None when no record
    matches the workload, precision and kernel."""
    path = os.path.join(REPO, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        table = json.load(f)
    key = f"{workload}/{precision}/{kernel}"
    rec = table.get(key)
    return (None, None) if rec is None else (rec["bytes_per_launch"], f"profiles/{rec['source']} ({key})")


def product_ms(prof):
    """Device milliseconds of the product kernels (names starting gemm) in one
    step, from the per-launch HIP event timings."""
    return sum(ms for k, (n, ms) in prof.items() if k.startswith("gemm"))


def mixed_vs_plain(ia, args):
    """The 8192^3 float32 product plain vs split-BF16 and the 6000^3 float64
    product plain vs the Ozaki forms (slice or modulus count autotuned to a
    1e-13 componentwise target): ms per step end to end, GFLOP/s, device time
    of the product kernels, dominant kernel and errors of each form."""
    out = {}
    for name, mixed in (("gemm8k_f32", ("split_bf16",)), ("gemm6k_f64", ("ozaki_crt", "ozaki_int8"))):
        w = make_workload(ia, name)
        rec = {}
        for prec in ("plain",) + mixed:
            pc = PRECISIONS[prec]
            step = lambda: ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"], precision=pc, slices=0)  # noqa: E731
            ms, c, prof = timed_steps(ia, step, args.mixed_warmup, args.mixed_steps)
            S = ia.last_ozaki_slices()[0] if prec.startswith("ozaki") else None
            roof = roofline(ia, w, prof, prec, npairs(prec, S))
            pms = product_ms(prof)
            rec[prec] = {"ms_per_step": round(ms, 3), "gflops": round(w["flops"] / ms / 1e6, 1),
                         "product_kernels_ms": round(pms, 3),
                         "product_gflops": round(w["flops"] / pms / 1e6, 1) if pms > 0 else None,
                         "dominant_kernel": roof["kernel"], "kernel_frac": roof["frac"],
                         "kernel_peak_unit": f'{roof["peak"]} {roof["unit"]}',
                         "errors_vs_fp64_oracle": error_metrics(c.to_numpy(), w["a_dense"], w["b_dense"])}
            if S:
                rec[prec]["ozaki_count"] = S
            print(f"# mixed_vs_plain {name} {prec}: {json.dumps(rec[prec])}", file=sys.stderr, flush=True)
        for m in mixed:
            rec[f"ratio_{m}_over_plain"] = {
                "end_to_end": round(rec[m]["gflops"] / rec["plain"]["gflops"], 3),
                "product_kernels": round(rec[m]["product_gflops"] / rec["plain"]["product_gflops"], 3)
                if rec[m]["product_gflops"] and rec["plain"]["product_gflops"] else None}
        out[name] = rec
        del w
    return out


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
    t_setup = time.time()
    w = make_workload(ia, args.workload)
    config = w["config"]
    if args.workload == "batch512":
        step = lambda: ia.matmul_batch(w["pairs"], w["c_chunks_list"], w["c_blocks_list"])  # noqa: E731
    else:
        prec = PRECISIONS[args.precision]
        config["precision"] = args.precision
        step = lambda: ia.matmul(w["a"], w["b"], w["c_chunks"], w["c_blocks"], precision=prec,  # noqa: E731
                                 slices=args.slices)
    print(f"# setup (generate + host encode) {time.time() - t_setup:.1f}s", file=sys.stderr, flush=True)
    nb, cb = w["a"].info()

    ms, c, prof = timed_steps(ia, step, args.warmup, args.steps)
    enc_s, _ = ia.last_matmul_times()
    S = None
    if args.precision.startswith("ozaki"):
        S, _ = ia.last_ozaki_slices()
        config["ozaki_count"] = S
        config["ozaki_count_chosen_by"] = "argument" if args.slices else "sampled-error autotune (target 1e-13)"
    print(f"# kernels: {json.dumps({k: [v[0], round(v[1] / max(v[0], 1), 4)] for k, v in prof.items()})}",
          file=sys.stderr, flush=True)
    if args.workload == "batch512":
        # the batch runs as graph replays: its kernels are timed on one more,
        # eager replay of the same launches
        os.environ["IARRAY_BATCH_EAGER"] = "1"
        ia.prof_enable(True)
        ia.prof_reset()
        step()
        ia.device_sync()
        prof = dict(ia.prof_entries())
        ia.prof_enable(False)
        del os.environ["IARRAY_BATCH_EAGER"]
        print(f"# kernels (eager replay): {json.dumps({k: [v[0], round(v[1] / max(v[0], 1), 4)] for k, v in prof.items()})}",
              file=sys.stderr, flush=True)
    roof = roofline(ia, w, prof, args.precision, npairs(args.precision, S))
    traffic, tsrc = profile_traffic(args.workload, args.precision, roof["kernel"])
    roof["traffic"] = traffic
    roof["traffic_source"] = tsrc or "not profiled for this kernel and workload"

    errors = None
    if args.workload != "batch512":
        errors = error_metrics(c.to_numpy(), w["a_dense"], w["b_dense"])
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(args, w) if args.workload == "batch512" else cpu_baseline(args, w)
    step_s = ms / 1e3
    out = {"metric": METRIC, "value": round(w["flops"] / step_s / 1e9, 3), "unit": "GFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": w["dtype"], "data": "synthetic (seeded numpy generator)",
           "config": config, "roofline": roof, "cpu_baseline": cpu, "errors_vs_fp64_oracle": errors,
           "compressed_bytes_A": cb, "decoded_bytes_A": nb,
           "host_encode_ms_last_step": round(enc_s * 1e3, 3),
           "compressed_GBps": round(cb / step_s / 1e9, 2), "decoded_GBps": round(nb / step_s / 1e9, 2)}
    if args.workload == "gemv20k":
        M, K = config["M"], config["K"]
        step_bytes = cb + 4 * (K + M)
        out["step_roofline"] = {"bound": "hbm", "algorithmic_bytes": step_bytes,
                                "achieved": round(step_bytes / step_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": round(step_bytes / step_s / 1e9 / HBM_PEAK_GBS, 4),
                                "note": "whole step (decode + GEMV + y write) against compressed bytes of A + 4(K+M)"}
    if args.workload == "gemv20k" and not args.no_mixed:
        del w, c
        out["mixed_vs_plain"] = mixed_vs_plain(ia, args)
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


def cpu_baseline(args, w):
    """The reference's CPU path (oracle restatement of gemm_blosc / gemv_blosc
    with block threads, liblz4 decode, single-threaded MKL per block, LZ4
    encode of C) on the whole workload once."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    blas = orc.blas_name() if orc.set_blas(1) else "plain-c"
    orc.set_lz4(1)
    ca, cb = w["a"], w["b"]
    a = orc.Array(ca.chunks(), ca.shape, *w["a_tiling"], ca.dtype.itemsize)
    b = orc.Array(cb.chunks(), cb.shape, *w["b_tiling"], cb.dtype.itemsize)
    cc, cbk = w["c_chunks"], w["c_blocks"]
    if len(cb.shape) == 1:
        fn = lambda: orc.gemv_blosc(a, b, cc, cbk, nthreads=threads)  # noqa: E731
        filt = 2
    else:
        fn = lambda: orc.gemm_blosc(a, b, cc, cbk, nthreads=threads)  # noqa: E731
        filt = 1
    t0 = time.perf_counter()
    c = fn()
    t_mm = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.encode_dense(np.ascontiguousarray(c), cc, cbk, 5, filt, threads)
    t_enc = time.perf_counter() - t0
    total = t_mm + t_enc
    return {"value": round(w["flops"] / total / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"the full workload once (matmul {t_mm:.2f}s + C encode {t_enc:.2f}s), blas={blas}, "
                      f"lz4=liblz4"}


def cpu_baseline_batch(args, w, nsample=48):
    """The reference runs the 512 products as sequential iarray_linalg_matmul
    calls; time the first `nsample` of them (decode + blocked GEMM + C encode,
    block-parallel over the host cores) and report their GFLOP/s."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    blas = orc.blas_name() if orc.set_blas(1) else "plain-c"
    orc.set_lz4(1)
    flops = 0.0
    t0 = time.perf_counter()
    for (ca, cb), cs, bs in list(zip(w["pairs"], w["c_chunks_list"], w["c_blocks_list"]))[:nsample]:
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


if __name__ == "__main__":
    main()
