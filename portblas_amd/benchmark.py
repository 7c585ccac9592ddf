"""portBLAS's benchmark front end, run on this library.

    python -m portblas_amd.benchmark gemm --csv-param=params.csv \\
        --benchmark_out=results.json --benchmark_out_format=json

The reference's benchmark suite (benchmark/README.md:111-166) reads one
parameter set per CSV line, runs each operator many times and reports Google
Benchmark counters.  This module keeps that interface for the operators of
the hot path and for SYMM and TRSM, so the reference's parameter files (benchmark/config_csv/...)
run unchanged and the results compare counter for counter:

* CSV formats and default ranges: common/include/common/common_utils.hpp
  (get_blas1_params :286-306, get_blas2_params :316-, get_blas3_params
  :469-517, get_gemm_batched_params :725-, get_gemm_batched_strided_params
  :787-, get_reduction_params :900-921, get_symm_params :933-966,
  get_trsm_params :1015-1049; empty
  lines are skipped, complex alpha/beta pairs keep their real parts);
* benchmark names: common/include/common/benchmark_names.hpp
  ("BM_<Op><type>/<params>/<mem_type>");
* counters: common/include/common/blas1_state_counters.hpp,
  blas2_state_counters.hpp:73-100, blas3_state_counters.hpp:33-74 and
  :139-167 (SYMM, benchmark/portblas/blas3/symm.cpp) and :226-253 (TRSM,
  benchmark/portblas/blas3/trsm.cpp), and
  benchmark/portblas/extension/reduction.cpp:41-49 (n_fl_ops,
  bytes_processed, the sizes), common_utils.hpp:1893-1919 (best / total /
  average event and overall times, in ns);
* warm-up: 10 calls (common_utils.hpp:1841-1845).

"overall" time is the host clock around one call including its
synchronisation (the reference's sb_handle.wait inside the timed lambda);
"event" time is the HIP-event time of the call on the handle's stream.
Memory type "usm": operands are device allocations.
"""
from __future__ import annotations

import argparse
import datetime
import json
import math
import re
import socket
import sys
import time
from dataclasses import dataclass, field
from typing import Callable

MEM_TYPE = "usm"
OPS = ("gemm", "gemm_batched", "gemm_batched_strided", "gemv", "dot", "asum", "nrm2", "iamax", "iamin",
       "reduction", "symm", "trsm")
OP_NAMES = {"gemm": "Gemm", "gemm_batched": "Gemm_batched", "gemm_batched_strided": "Gemm_batched_strided",
            "gemv": "Gemv", "dot": "Dot", "asum": "Asum", "nrm2": "Nrm2", "iamax": "Iamax", "iamin": "Iamin",
            "reduction": "Reduction", "symm": "Symm", "trsm": "Trsm"}
ELEM_SIZE = {"float": 4, "double": 8}


class CsvError(ValueError):
    """A parameter file line that does not parse (the reference prints the
    line and exits)."""


# ---------------------------------------------------------------------------
# CSV parameter files
# ---------------------------------------------------------------------------
def _cell(s: str) -> str:
    return s.replace("\0", "").strip()


def _int(s: str) -> int:
    return int(_cell(s))


def _scalar(s: str) -> float:
    return float(_cell(s))


def _trans(s: str) -> str:
    t = _cell(s)
    if t not in ("n", "t", "c", "N", "T", "C"):
        raise ValueError(f"invalid transpose '{t}'")
    return t


def _batch_type(s: str) -> int:
    t = _cell(s)
    if t == "strided":
        return 0
    if t == "interleaved":
        return 1
    raise ValueError(f"Unrecognized batch type: '{t}'")


def _batch_type_str(b: int) -> str:
    return {0: "strided", 1: "interleaved"}.get(b, "unknown")


def parse_csv(path: str, op: str) -> list:
    """Parameter tuples of `op` from a CSV file, one per non-empty line."""
    parse = _PARSERS[op]
    out = []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\r\n")
            if not line:
                continue
            cells = line.split(",")
            try:
                out.append(parse(cells))
            except Exception as e:  # noqa: BLE001 - the reference reports any parse error the same way
                raise CsvError(f"Error while parsing CSV file on line '{line}': {e}") from e
    if not out:
        raise CsvError("No data has been read. The given file might not exist or be empty.")
    return out


def _p_blas1(v):
    if len(v) != 1:
        raise ValueError("invalid number of parameters (1 expected)")
    return (_int(v[0]),)


def _p_gemv(v):
    if len(v) == 5:
        return (_trans(v[0]), _int(v[1]), _int(v[2]), _scalar(v[3]), _scalar(v[4]))
    if len(v) == 7:  # complex alpha, beta: real parts
        return (_trans(v[0]), _int(v[1]), _int(v[2]), _scalar(v[3]), _scalar(v[5]))
    raise ValueError("invalid number of parameters (5 or 7 expected)")


def _p_gemm(v):
    if len(v) == 7:
        return (_trans(v[0]), _trans(v[1]), _int(v[2]), _int(v[3]), _int(v[4]), _scalar(v[5]), _scalar(v[6]))
    if len(v) == 9:
        return (_trans(v[0]), _trans(v[1]), _int(v[2]), _int(v[3]), _int(v[4]), _scalar(v[5]), _scalar(v[7]))
    raise ValueError("invalid number of parameters (7 or 9 expected)")


def _p_gemm_batched(v):
    if len(v) == 9:
        return (_trans(v[0]), _trans(v[1]), _int(v[2]), _int(v[3]), _int(v[4]), _scalar(v[5]), _scalar(v[6]),
                _int(v[7]), _batch_type(v[8]))
    if len(v) == 11:
        return (_trans(v[0]), _trans(v[1]), _int(v[2]), _int(v[3]), _int(v[4]), _scalar(v[5]), _scalar(v[7]),
                _int(v[9]), _batch_type(v[10]))
    raise ValueError("invalid number of parameters (9 or 11 expected)")


def _p_gemm_batched_strided(v):
    if len(v) == 11:
        return (_trans(v[0]), _trans(v[1]), _int(v[2]), _int(v[3]), _int(v[4]), _scalar(v[5]), _scalar(v[6]),
                _int(v[7]), _int(v[8]), _int(v[9]), _int(v[10]))
    if len(v) == 13:
        return (_trans(v[0]), _trans(v[1]), _int(v[2]), _int(v[3]), _int(v[4]), _scalar(v[5]), _scalar(v[7]),
                _int(v[9]), _int(v[10]), _int(v[11]), _int(v[12]))
    raise ValueError("invalid number of parameters (11 or 13 expected)")


def _p_reduction(v):
    if len(v) != 2:
        raise ValueError("invalid number of parameters (2 expected)")
    return (_int(v[0]), _int(v[1]))


def _p_symm(v):
    """side, uplo (first character of each), m, n, alpha, beta
    (common_utils.hpp:949-964)."""
    if len(v) != 6:
        raise ValueError("invalid number of parameters (6 expected)")
    flags = [_cell(x)[:1] for x in v[:2]]
    if not all(flags):
        raise ValueError("invalid parameter")
    return (*flags, _int(v[2]), _int(v[3]), _scalar(v[4]), _scalar(v[5]))


def _p_trsm(v):
    """side, uplo, trans, diag (first character of each), m, n, alpha
    (common_utils.hpp:1033-1047)."""
    if len(v) != 7:
        raise ValueError("invalid number of parameters (7 expected)")
    flags = [_cell(x)[:1] for x in v[:4]]
    if not all(flags):
        raise ValueError("invalid parameter")
    return (*flags, _int(v[4]), _int(v[5]), _scalar(v[6]))


_PARSERS = {"gemm": _p_gemm, "gemm_batched": _p_gemm_batched, "gemm_batched_strided": _p_gemm_batched_strided,
            "gemv": _p_gemv, "dot": _p_blas1, "asum": _p_blas1, "nrm2": _p_blas1, "iamax": _p_blas1,
            "iamin": _p_blas1, "reduction": _p_reduction, "symm": _p_symm,
            "trsm": _p_trsm}


def default_params(op: str) -> list:
    """The reference's ranges when no CSV file is given."""
    if op in ("dot", "asum", "nrm2", "iamax", "iamin"):
        return [(1 << e,) for e in range(10, 23)]
    if op == "gemv":
        return [(t, n, n, 1.0, 1.0) for t in ("n", "t") for n in (32, 128, 512, 2048, 8192)]
    if op == "gemm":
        d = (32, 256, 2048)
        return [(t1, t2, m, k, n, 1.0, 1.0) for t1 in ("n", "t") for t2 in ("n", "t")
                for m in d for k in d for n in d]
    if op == "gemm_batched":
        return [(t1, t2, n, n, n, 1.0, 1.0, 8, 0) for t1 in ("n", "t") for t2 in ("n", "t")
                for n in (128, 1024, 8192)]
    if op == "gemm_batched_strided":
        return [(t1, t2, m, m, m, 1.0, 1.0, 8, 2, 2, 2) for t1 in ("n", "t") for t2 in ("n", "t")
                for m in (128, 1024, 8192)]
    if op == "reduction":
        return [(r, r) for r in (256, 512, 1024, 2048, 4096, 8192)]
    if op == "symm":  # common_utils.hpp:934-948: m = n from 32 to 8192 by 4x, alpha = beta = 1
        return [(s, u, m, m, 1.0, 1.0) for s in ("l", "r") for u in ("u", "l") for m in (32, 128, 512, 2048, 8192)]
    if op == "trsm":  # common_utils.hpp:1016-1032: m = n from 512 to 8192 by 4x
        return [(s, u, t, d, m, m, 1.0) for s in ("l", "r") for u in ("u", "l") for t in ("n", "t")
                for d in ("u", "n") for m in (512, 2048, 8192)]
    raise ValueError(f"unknown operator '{op}'")


# ---------------------------------------------------------------------------
# Names and counters
# ---------------------------------------------------------------------------
def _fmt(v) -> str:
    if isinstance(v, float):
        return f"{v:g}"
    return str(v)


def benchmark_names(op: str, dtype: str, p: tuple) -> list:
    """Names of the benchmark(s) one parameter tuple registers (reduction
    registers an inner and an outer one)."""
    base = f"BM_{OP_NAMES[op]}<{dtype}>"
    if op in ("dot", "asum", "nrm2", "iamax", "iamin"):
        args = [p[0]]
    elif op == "gemv":
        args = [p[0], p[1], p[2]]
    elif op == "gemm":
        args = list(p[:5])
    elif op == "gemm_batched":
        args = list(p[:5]) + [p[7], _batch_type_str(p[8])]
    elif op == "gemm_batched_strided":
        args = list(p[:5]) + list(p[7:11])
    elif op in ("symm", "trsm"):
        args = list(p[:6])
    else:  # reduction
        return [f"{base}/{p[0]}/{p[1]}/{d}/{MEM_TYPE}" for d in ("inner", "outer")]
    return [f"{base}/" + "/".join(_fmt(a) for a in args) + f"/{MEM_TYPE}"]


def counters(op: str, dtype: str, p: tuple) -> dict:
    """The reference's state counters for one parameter tuple."""
    es = ELEM_SIZE[dtype]
    if op in ("dot", "asum", "nrm2", "iamax", "iamin"):
        n = float(p[0])
        if op == "dot":
            return {"size": n, "n_fl_ops": 2 * n, "bytes_processed": (2 * n + 1) * es}
        if op == "nrm2":
            return {"size": n, "n_fl_ops": 2 * n, "bytes_processed": (n + 1) * es}
        return {"size": n, "n_fl_ops": n, "bytes_processed": (n + 1) * es}
    if op == "gemv":
        t, m, n, _, beta = p
        m, n = float(m), float(n)
        xlen, ylen = (n, m) if t == "n" else (m, n)
        flops = 2.0 * m * n + xlen + (2 * ylen if beta != 0 else 0)
        nbytes = (m * n + xlen + ylen + (ylen if beta != 0 else 0)) * es
        return {"beta": float(beta), "m": m, "n": n, "n_fl_ops": flops, "bytes_processed": nbytes}
    if op in ("gemm", "gemm_batched", "gemm_batched_strided"):
        m, k, n, beta = float(p[2]), float(p[3]), float(p[4]), p[6]
        batch = float(p[7]) if op != "gemm" else 1.0
        c = {"beta": float(beta), "m": m, "n": n, "k": k, "batch_size": batch}
        if op == "gemm_batched_strided":
            c.update(stride_a_mul=float(p[8]), stride_b_mul=float(p[9]), stride_c_mul=float(p[10]))
        flops = (2 * k * m * n + m * n + (2 * m * n if beta != 0 else 0)) * batch
        nbytes = (m * k + k * n + m * n + (m * n if beta != 0 else 0)) * batch * es
        c.update(n_fl_ops=flops, bytes_processed=nbytes)
        return c
    if op == "symm":
        # blas3_state_counters.hpp:139-167, kept as written there (for side
        # 'r' the product counts 2 n^3 flops).
        side, m, n, beta = p[0], float(p[2]), float(p[3]), p[5]
        flops = (2 * m * m * n if side == "l" else 2 * n * n * n) + (2 * m * n if beta != 0 else 0)
        read_a = m * (m + 1) / 2 if side == "l" else n * (n + 1) / 2
        nbytes = ((2 if beta != 0 else 1) * m * n + m * n + read_a) * es
        return {"beta": float(beta), "m": m, "n": n, "n_fl_ops": flops, "bytes_processed": nbytes}
    if op == "trsm":
        # blas3_state_counters.hpp:226-253, kept as written there (for side
        # 'r' the A term counts m n elements and the B term n (n + 1) / 2).
        side, m, n = p[0], float(p[4]), float(p[5])
        tri = (m if side == "l" else n)
        flops = 2 * tri * (tri + 1) / 2 * (n if side == "l" else m) + m * n
        read_a = m * (m + 1) / 2 if side == "l" else m * n
        read_b = n * (n + 1) / 2 if side == "r" else m * n
        return {"beta": 0.0, "m": m, "n": n, "n_fl_ops": flops, "bytes_processed": (read_a + read_b + m * n) * es}
    rows, cols = float(p[0]), float(p[1])
    return {"rows": rows, "cols": cols, "n_fl_ops": rows * cols, "bytes_processed": rows * cols * es}


# ---------------------------------------------------------------------------
# Running
# ---------------------------------------------------------------------------
@dataclass
class Result:
    name: str
    iterations: int
    counters: dict = field(default_factory=dict)


def _operator(op: str, dtype: str, p: tuple, h, torch, variant: str | None = None) -> Callable[[], None]:
    """Allocates the operands of one parameter tuple and returns a callable
    that runs the operator once and waits for it (as the reference's timed
    lambdas do)."""
    import portblas_amd as pb

    dt = {"float": torch.float32, "double": torch.float64}[dtype]
    gen = torch.Generator(device="cuda").manual_seed(0)
    rnd = lambda n: torch.rand(max(int(n), 1), generator=gen, device="cuda", dtype=dt) * 2 - 1  # noqa: E731
    if op in ("dot", "asum", "nrm2", "iamax", "iamin"):
        n = p[0]
        x = rnd(n) / max(n, 1)
        y = rnd(n)
        if op in ("iamax", "iamin"):
            rs = torch.zeros(1, dtype=torch.int32, device="cuda")
        else:
            rs = torch.zeros(1, dtype=dt, device="cuda")
        call = {"dot": lambda: pb._dot(h, n, x, 1, y, 1, rs), "asum": lambda: pb._asum(h, n, x, 1, rs),
                "nrm2": lambda: pb._nrm2(h, n, x, 1, rs), "iamax": lambda: pb._iamax(h, n, x, 1, rs),
                "iamin": lambda: pb._iamin(h, n, x, 1, rs)}[op]
    elif op == "gemv":
        t, m, n, alpha, beta = p
        xlen, ylen = (n, m) if t in "nN" else (m, n)
        A, x, y = rnd(m * n), rnd(xlen), rnd(ylen)
        call = lambda: pb._gemv(h, t, m, n, alpha, A, max(m, 1), x, 1, beta, y, 1)  # noqa: E731
    elif op in ("gemm", "gemm_batched", "gemm_batched_strided"):
        ta, tb, m, k, n, alpha, beta = p[:7]
        lda = m if ta in "nN" else k
        ldb = k if tb in "nN" else n
        sa, sb, sc = m * k, k * n, m * n
        if op == "gemm":
            A, B, C = rnd(sa), rnd(sb), rnd(sc)
            call = lambda: pb._gemm(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, m)  # noqa: E731
        elif op == "gemm_batched":
            batch, btype = p[7], p[8]
            A, B, C = rnd(sa * batch), rnd(sb * batch), rnd(sc * batch)
            call = lambda: pb._gemm_batched(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, m,  # noqa: E731
                                            batch, pb.gemm_batch_type_t(btype))
        else:
            batch, ma, mb, mc = p[7:11]
            A, B, C = rnd(sa * ma * batch), rnd(sb * mb * batch), rnd(sc * mc * batch)
            call = lambda: pb._gemm_strided_batched(h, ta, tb, m, n, k, alpha, A, lda, sa * ma, B, ldb,  # noqa: E731
                                                    sb * mb, beta, C, m, sc * mc, batch)
    elif op == "symm":
        side, uplo, m, n, alpha, beta = p
        k = m if side in "lL" else n
        A, B, C = rnd(k * k), rnd(m * n), rnd(m * n)
        call = lambda: pb._symm(h, side, uplo, m, n, alpha, A, k, B, m, beta, C, m)  # noqa: E731
    elif op == "trsm":
        side, uplo, trans, diag, m, n, alpha = p
        k = m if side in "lL" else n
        # Triangle uniform in +-d / k and diagonal d (d in [1, 10] unless unit):
        # diagonally dominant, as the reference's fill_trsm_matrix
        # (common_utils.hpp:1758-1781) makes it; the time does not depend on
        # the values.
        d = 1.0 if diag in "uU" else 1.0 + 9.0 * float(torch.rand(1, generator=torch.Generator().manual_seed(0)))
        R = rnd(k * k).view(k, k) * (d / k)  # R[j][i] holds A(i, j): the column-major matrix
        T = torch.triu(R) if uplo in "lL" else torch.tril(R)
        T.diagonal().fill_(d)
        A = T.reshape(-1)
        B = rnd(m * n)
        call = lambda: pb._trsm(h, side, uplo, trans, diag, m, n, alpha, A, k, B, m)  # noqa: E731
    else:
        rows, cols = p
        dim = pb.reduction_dim_t[variant]
        mat = rnd(rows * cols)
        out = rnd(rows if dim == pb.reduction_dim_t.outer else cols)
        call = lambda: pb.extension._reduction(pb.AddOperator, h, mat, rows, out, rows, cols, dim)  # noqa: E731

    def run():
        call()
        h.wait()
    return run


def run_one(fn: Callable[[], None], torch, min_time: float, max_iters: int) -> dict:
    """Warm-up (10 calls), then timed iterations until `min_time` seconds or
    `max_iters`; returns the reference's time counters (ns)."""
    for _ in range(10):
        fn()
    torch.cuda.synchronize()
    best_e = best_o = math.inf
    tot_e = tot_o = 0.0
    it = 0
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    start = time.perf_counter()
    while it < max_iters and (it < 1 or time.perf_counter() - start < min_time):
        t0 = time.perf_counter_ns()
        e0.record()
        fn()
        e1.record()
        t1 = time.perf_counter_ns()
        e1.synchronize()
        ev = e0.elapsed_time(e1) * 1e6
        ov = float(t1 - t0)
        tot_e += ev
        tot_o += ov
        best_e, best_o = min(best_e, ev), min(best_o, ov)
        it += 1
    return {"iterations": it, "best_event_time": best_e, "best_overall_time": best_o,
            "total_event_time": tot_e, "total_overall_time": tot_o, "avg_event_time": tot_e / it,
            "avg_overall_time": tot_o / it}


def run(op: str, params: list, dtype: str = "float", min_time: float = 0.5, max_iters: int = 100000,
        name_filter: str | None = None, out=sys.stdout) -> list:
    """Runs every parameter tuple of `op`; returns Google-Benchmark style
    records (dicts)."""
    import torch

    import portblas_amd as pb

    if not torch.cuda.is_available():
        raise RuntimeError("the benchmarks need a GPU")
    h = pb.SB_Handle()
    rx = re.compile(name_filter) if name_filter else None
    records = []
    for p in params:
        names = benchmark_names(op, dtype, p)
        variants = ["inner", "outer"] if op == "reduction" else [None]
        for name, variant in zip(names, variants):
            if rx and not rx.search(name):
                continue
            fn = _operator(op, dtype, p, h, torch, variant)
            t = run_one(fn, torch, min_time, max_iters)
            c = counters(op, dtype, p)
            it = t.pop("iterations")
            real = t["avg_overall_time"]
            rec = {"name": name, "family_index": len(records), "per_family_instance_index": 0, "run_name": name,
                   "run_type": "iteration", "repetitions": 1, "repetition_index": 0, "threads": 1,
                   "iterations": it, "real_time": real, "cpu_time": real, "time_unit": "ns",
                   "bytes_per_second": c["bytes_processed"] / (real * 1e-9),
                   "items_per_second": c["n_fl_ops"] / (real * 1e-9)}
            rec.update(c)
            rec.update(t)
            records.append(rec)
            print(f"{name:<60s} {real:14.0f} ns {it:10d}  "
                  f"{c['n_fl_ops'] / t['avg_event_time']:9.2f} GFLOP/s  "
                  f"{c['bytes_processed'] / t['avg_event_time']:9.2f} GB/s (event)", file=out, flush=True)
            del fn
    return records


def context(torch, h=None) -> dict:
    import portblas_amd as pb

    ctx = {"date": datetime.datetime.now().astimezone().isoformat(timespec="seconds"),
           "host_name": socket.gethostname(), "executable": "python -m portblas_amd.benchmark",
           "library": "portblas_amd", "library_build_type": "release"}
    try:
        ctx["device"] = torch.cuda.get_device_name(0)
        ctx["compute_units"] = (h or pb.SB_Handle()).get_num_compute_units()
    except Exception:  # noqa: BLE001 - context is informational
        pass
    return ctx


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("operator", choices=OPS)
    ap.add_argument("--csv-param", "--csv_param", dest="csv_param", default="")
    ap.add_argument("--type", dest="dtype", choices=("float", "double"), default="float")
    ap.add_argument("--benchmark_out", default="")
    ap.add_argument("--benchmark_out_format", choices=("json", "csv", "console"), default="json")
    ap.add_argument("--benchmark_filter", default=None)
    ap.add_argument("--benchmark_min_time", type=float, default=0.5, help="seconds per benchmark")
    ap.add_argument("--max-iterations", type=int, default=100000)
    args = ap.parse_args(argv)
    if args.csv_param:
        params = parse_csv(args.csv_param, args.operator)
    else:
        print("WARNING: no CSV parameter file has been given. Default ranges will be used.", file=sys.stderr)
        params = default_params(args.operator)
    import torch

    records = run(args.operator, params, args.dtype, args.benchmark_min_time, args.max_iterations,
                  args.benchmark_filter)
    if args.benchmark_out:
        if args.benchmark_out_format == "csv":
            keys = list(dict.fromkeys(k for r in records for k in r))
            with open(args.benchmark_out, "w") as f:
                f.write(",".join(keys) + "\n")
                for r in records:
                    f.write(",".join(json.dumps(r.get(k, "")) for k in keys) + "\n")
        else:
            with open(args.benchmark_out, "w") as f:
                json.dump({"context": context(torch), "benchmarks": records}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
