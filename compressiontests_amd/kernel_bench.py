"""Kernel-level measurement of the SpMV and smoother kernels (config 2: 64^3 cantilever,
element moduli 10^U(-2,2), FP64 values / int32 indices, x ~ N(0,1)): average device time per
launch over `--repeats` back-to-back launches (HIP events), algorithmic bytes per launch and
the resulting GB/s against the 8 TB/s HBM peak.

  python -m compressiontests_amd.kernel_bench --n 64 --kinds spmv,jacobi,residual,spmv_f32,copy

`copy` is a plain device-to-device copy kernel of known size, the calibration for rocprofv3's
FETCH_SIZE / WRITE_SIZE counters; `read` a streaming read of the same byte count (16-byte
non-temporal loads, nothing written), the practical HBM read ceiling.  `spmm64_k32` / `spmm32_k32` (any k in 4, 8, 16, 32): the
block SpMM Y = A X of the batching form, FP64 or FP32 values and vectors (suffix `_node`: the
node-triple CSR kernel; `_nb`: the node-block kernel the block solver uses on the level
matrices, on a 3x3 block copy built from the CSR; suffix `_dot`: the block-PCG form q = A p with the
per-column p.q partials).  `duvi` / `duvi_u32` / `duvi_raw` (DuVi block format: deduplicated
values with the narrowest index, forced 32-bit indices, raw values): the kernel the library
chooses; suffix `_wave` forces the one-wavefront-per-row form.  `--moduli uniform` gives the
homogeneous cantilever, whose few distinct values take one-byte indices.
"""
import argparse
import ctypes as C
import json

import numpy as np

from . import _lib as L
from .device import DeviceArray, DeviceDuvi, cantilever_matrix, duvi_encode, rng_normal, rng_uniform
from . import kernels

HBM_PEAK_GBS = 8000.0


def spmv_bytes(nnz, n, val_bytes=8, vec_bytes=8):
    return (4 + val_bytes) * nnz + 4 * (n + 1) + 2 * vec_bytes * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--kinds", default="spmv,jacobi,residual,spmv_f32,copy")
    ap.add_argument("--repeats", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=20261203 + 2)
    ap.add_argument("--moduli", default="loguniform", choices=["loguniform", "lognormal", "uniform"])
    args = ap.parse_args()
    n = args.n
    if args.moduli == "loguniform":
        dens = 10 ** (4 * rng_uniform(args.seed, n ** 3) - 2)
    elif args.moduli == "lognormal":   # log-normal, sigma 1.5, clipped to a 1e3 contrast
        dens = np.exp(np.clip(1.5 * rng_normal(20261203 + 3, n ** 3), -0.5 * np.log(1e3), 0.5 * np.log(1e3)))
    else:   # one modulus everywhere: few distinct values, so DuVi's value indices fit one byte
        dens = None
    A = cantilever_matrix(n, n, n, dens=dens)
    N, nnz = A.nrows, A.nnz
    x = DeviceArray.from_numpy(rng_normal(args.seed + 1, N))
    y = DeviceArray(N)
    b = DeviceArray.from_numpy(rng_normal(args.seed + 2, N))
    pre = DeviceArray.from_numpy(np.full(N, 0.6 / 100.0))
    out = {"n": n, "rows": N, "nnz": nnz, "repeats": args.repeats, "kernels": {}}
    for kind in args.kinds.split(","):
        if kind == "spmv":
            ms = kernels.time_csr_kernel(0, A, x, y, warmup=args.warmup, repeats=args.repeats)
            by = spmv_bytes(nnz, N)
        elif kind == "jacobi":
            ms = kernels.time_csr_kernel(1, A, x, y, b=b, pre=pre, warmup=args.warmup, repeats=args.repeats)
            by = spmv_bytes(nnz, N) + 16 * N
        elif kind == "spmv_dot":
            ms = kernels.time_csr_kernel(4, A, x, y, warmup=args.warmup, repeats=args.repeats)
            by = spmv_bytes(nnz, N)
        elif kind == "residual":
            ms = kernels.time_csr_kernel(2, A, x, y, b=b, warmup=args.warmup, repeats=args.repeats)
            by = spmv_bytes(nnz, N) + 8 * N
        elif kind == "spmv_f32":
            h = A.to_host()
            from .device import DeviceCsr
            A32 = DeviceCsr.arrays_from_host(h, val_dtype=np.float32)
            x32 = DeviceArray.from_numpy(x.numpy().astype(np.float32))
            y32 = DeviceArray(N, np.float32)
            ms = kernels.time_csr_kernel(3, A32, x32, y32, warmup=args.warmup, repeats=args.repeats)
            by = spmv_bytes(nnz, N, 4, 4)
        elif kind.startswith("spmm"):
            fp32 = kind.startswith("spmm32")
            node = "_node" in kind
            nodeblock = "_nb" in kind
            dot = kind.endswith("_dot")
            k = int(kind.split("_k")[1].split("_")[0])
            dt = np.float32 if fp32 else np.float64
            from .device import DeviceCsr
            Ak = DeviceCsr.arrays_from_host(A.to_host(), val_dtype=dt) if fp32 else A
            X = DeviceArray.from_numpy(rng_normal(args.seed + 3, N * k).astype(dt))
            Y = DeviceArray(N * k, dt)
            ms = kernels.time_spmm(Ak, X, Y, k, warmup=args.warmup, repeats=args.repeats, node=node, dot=dot,
                                   nodeblock=nodeblock)
            vb = 4 if fp32 else 8
            if nodeblock:   # 3x3 blocks: one column index and one block offset per 9 values / node
                by = vb * nnz + 4 * (nnz // 9) + 4 * (N // 3 + 1) + 2 * vb * k * N
            else:
                by = (vb + 4) * nnz + 4 * (N + 1) + 2 * vb * k * N
        elif kind.startswith("duvi"):
            wave = kind.endswith("_wave")
            mode = {"duvi": 0, "duvi_u32": 4, "duvi_raw": 8}[kind[:-5] if wave else kind]
            d = duvi_encode(A.to_host(), 1e-10, mode)
            D = DeviceDuvi(d)
            ms = kernels.time_duvi_spmv(D, x, y, form=1 if wave else 0, warmup=args.warmup, repeats=args.repeats)
            s_vi = d["vi_bytes"]
            by = nnz * (1 + (s_vi if s_vi else 8)) + 4 * d["nblocks"] + 8 * (N + 1) + (8 * d["nuniq"] if s_vi else 0) \
                + 16 * N
        elif kind == "copy":
            src = DeviceArray(nnz)
            dst = DeviceArray(nnz)
            res = C.c_double()
            L.call("mgx_time_copy_kernel", src.vp, dst.vp, C.c_uint64(8 * nnz), args.warmup, args.repeats, C.byref(res))
            ms = res.value
            by = 16 * nnz
        elif kind == "read":
            src = DeviceArray(2 * nnz)
            sink = DeviceArray(2)
            res = C.c_double()
            L.call("mgx_time_read_kernel", src.vp, C.c_uint64(16 * nnz), sink.vp, args.warmup, args.repeats,
                   C.byref(res))
            ms = res.value
            by = 16 * nnz
        else:
            raise SystemExit("unknown kind " + kind)
        gbs = by / (ms * 1e-3) / 1e9
        out["kernels"][kind] = {"avg_ms": round(ms, 5), "bytes": int(by), "GBps": round(gbs, 1),
                                "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
        print(f"{kind:12s} {ms:9.4f} ms  {by / 1e6:9.1f} MB  {gbs:8.1f} GB/s  {gbs / HBM_PEAK_GBS:6.3f} of peak",
              flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
