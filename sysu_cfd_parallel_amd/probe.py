#!/usr/bin/env python3
"""SpMV counter probe: the channel workload's pressure-correction matrix
(assembled by one SIMPLE step from rest), then `repeats` launches of the
production SpMV kernel in FP64 and FP32, then one plain device copy of a known
size (k_stream_copy) whose FETCH_SIZE / WRITE_SIZE readings scale the
SpMV's counters to bytes.

Run under `rocprofv3 --pmc FETCH_SIZE` (then `--pmc WRITE_SIZE`) to measure the
HBM traffic of one SpMV launch; the printed JSON gives the algorithmic bytes
per launch and the copy size to compare against.

    python3 sysu_cfd_parallel_amd/probe.py --size full --repeats 20
"""
from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", choices=["full", "medium", "small"], default="full")
    ap.add_argument("--repeats", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=1024)
    args = ap.parse_args()

    import sysu_cfd_parallel_amd as s

    dims = {"full": (200, 100, 200), "medium": (100, 50, 100), "small": (40, 20, 40)}[args.size]
    case = s.channel(*dims, lx=2.0, ly=1.0, lz=2.0, jitter=0.15, split=0.10, seed=0, re=1000.0)
    p = s.set_all_ksp(s.default_params(), s.KSP_GMRES, 50, s.PC_JACOBI)
    for k in range(3):
        p.ksp[k].maxit = 20000
    p.precision = s.MIXED
    p.autotune = 1  # the production kernel shapes (SELL-64 SpMV on this mesh)
    copy_bytes = args.copy_mib << 20
    with s.Solver(case, p) as dev:
        z = dev.sizes
        dev.iterate(1)
        ms64 = dev.spmv_time(1, 0, args.repeats)
        ms32 = dev.spmv_time(1, 1, args.repeats)
        shapes = dev.kernel_shapes()
        ms_copy = dev.copy_time(copy_bytes, args.repeats)

    def spmv_bytes(bv, bx):
        # CSR values + column ids, row pointers, x, y and the Jacobi diagonal
        return (bv + 4) * z.nnz + 4 * (z.n_cell + 1) + 3 * bx * z.n_cell

    print(json.dumps({
        "cells": z.n_cell, "nnz": z.nnz, "repeats": args.repeats, "shapes": shapes,
        "spmv_fp64": {"bytes_per_launch": spmv_bytes(8, 8), "ms": ms64},
        "spmv_fp32": {"bytes_per_launch": spmv_bytes(4, 4), "ms": ms32},
        "copy": {"bytes_read": copy_bytes, "bytes_written": copy_bytes, "ms": ms_copy,
                 "gbs": 2 * copy_bytes / (ms_copy * 1e-3) / 1e9},
    }), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
