"""The portBLAS benchmark front end (portblas_amd.benchmark): CSV parameter
files in the reference's formats (benchmark/README.md:111-166), its default
ranges, benchmark names and counters — on the CPU — and one run of every
operator on the GPU (SYMM and TRSM included: benchmark/portblas/blas3/symm.cpp,
trsm.cpp)."""
from __future__ import annotations

import json

import pytest

from portblas_amd import benchmark as bm


def _csv(tmp_path, text, name="p.csv"):
    f = tmp_path / name
    f.write_text(text)
    return str(f)


def test_gemm_csv_scalar_and_complex_forms(tmp_path):
    # the README's two examples: 7 fields, and 9 with complex alpha/beta (real parts kept)
    p = bm.parse_csv(_csv(tmp_path, "n,n,42,42,42,1,0\nn,t,64,128,64,0.5,0.5\n\nt,n,13,3,7,0,0.7\n"), "gemm")
    assert p == [("n", "n", 42, 42, 42, 1.0, 0.0), ("n", "t", 64, 128, 64, 0.5, 0.5), ("t", "n", 13, 3, 7, 0.0, 0.7)]
    q = bm.parse_csv(_csv(tmp_path, "n,n,42,42,42,1,2,3,0\nn,t,64,128,64,0.5,1,0.5,1\n"), "gemm")
    assert q == [("n", "n", 42, 42, 42, 1.0, 3.0), ("n", "t", 64, 128, 64, 0.5, 0.5)]


def test_other_formats(tmp_path):
    assert bm.parse_csv(_csv(tmp_path, "1024\n2048\n"), "dot") == [(1024,), (2048,)]
    assert bm.parse_csv(_csv(tmp_path, "t,32,128,1,0\n"), "gemv") == [("t", 32, 128, 1.0, 0.0)]
    assert bm.parse_csv(_csv(tmp_path, "n,n,256,256,2,1,0,32,1,1,1\n"), "gemm_batched_strided") == [
        ("n", "n", 256, 256, 2, 1.0, 0.0, 32, 1, 1, 1)]
    assert bm.parse_csv(_csv(tmp_path, "t,n,64,64,64,1,1,8,interleaved\n"), "gemm_batched") == [
        ("t", "n", 64, 64, 64, 1.0, 1.0, 8, 1)]
    assert bm.parse_csv(_csv(tmp_path, "513,8195\n"), "reduction") == [(513, 8195)]
    assert bm.parse_csv(_csv(tmp_path, "l,u,n,n,512,256,1.5\nr,l,t,u,64,32,1\n"), "trsm") == [
        ("l", "u", "n", "n", 512, 256, 1.5), ("r", "l", "t", "u", 64, 32, 1.0)]
    assert bm.parse_csv(_csv(tmp_path, "r,l,64,32,1.5,0\n"), "symm") == [("r", "l", 64, 32, 1.5, 0.0)]


@pytest.mark.parametrize("op,text", [("gemm", "n,n,42,42\n"), ("gemm", "x,n,4,4,4,1,0\n"), ("dot", "12,3\n"),
                                     ("gemm_batched", "n,n,4,4,4,1,0,2,diagonal\n"), ("reduction", ""),
                                     ("trsm", "l,u,n,n,512,512\n"), ("trsm", "l,,n,n,4,4,1\n"),
                                     ("symm", "l,u,4,4,1\n")])
def test_bad_lines_are_reported(tmp_path, op, text):
    with pytest.raises(bm.CsvError):
        bm.parse_csv(_csv(tmp_path, text), op)


def test_default_ranges_match_the_reference():
    assert [p[0] for p in bm.default_params("asum")] == [2 ** e for e in range(10, 23)]
    g = bm.default_params("gemm")
    assert len(g) == 4 * 27 and {p[2] for p in g} == {32, 256, 2048} and g[0][5:] == (1.0, 1.0)
    assert [p[1] for p in bm.default_params("gemv") if p[0] == "n"] == [32, 128, 512, 2048, 8192]
    assert {p[2] for p in bm.default_params("gemm_batched_strided")} == {128, 1024, 8192}
    assert bm.default_params("gemm_batched_strided")[0][7:] == (8, 2, 2, 2)
    assert [p[0] for p in bm.default_params("reduction")] == [256, 512, 1024, 2048, 4096, 8192]
    t = bm.default_params("trsm")
    assert len(t) == 16 * 3 and {p[4] for p in t} == {512, 2048, 8192} and all(p[4] == p[5] for p in t)
    assert t[0] == ("l", "u", "n", "u", 512, 512, 1.0)
    y = bm.default_params("symm")
    assert len(y) == 4 * 5 and [p[2] for p in y[:5]] == [32, 128, 512, 2048, 8192] and y[0][4:] == (1.0, 1.0)


def test_names():
    assert bm.benchmark_names("gemm", "float", ("n", "t", 42, 43, 44, 1.0, 0.0)) == [
        "BM_Gemm<float>/n/t/42/43/44/usm"]
    assert bm.benchmark_names("dot", "double", (1024,)) == ["BM_Dot<double>/1024/usm"]
    assert bm.benchmark_names("gemv", "float", ("t", 32, 128, 1.0, 0.0)) == ["BM_Gemv<float>/t/32/128/usm"]
    assert bm.benchmark_names("gemm_batched", "float", ("n", "n", 8, 8, 8, 1.0, 1.0, 4, 1)) == [
        "BM_Gemm_batched<float>/n/n/8/8/8/4/interleaved/usm"]
    assert bm.benchmark_names("gemm_batched_strided", "float", ("n", "n", 8, 8, 8, 1.0, 1.0, 4, 2, 3, 1)) == [
        "BM_Gemm_batched_strided<float>/n/n/8/8/8/4/2/3/1/usm"]
    assert bm.benchmark_names("reduction", "float", (64, 32)) == [
        "BM_Reduction<float>/64/32/inner/usm", "BM_Reduction<float>/64/32/outer/usm"]
    assert bm.benchmark_names("trsm", "float", ("l", "u", "n", "u", 512, 256, 1.0)) == [
        "BM_Trsm<float>/l/u/n/u/512/256/usm"]
    assert bm.benchmark_names("symm", "double", ("r", "l", 32, 64, 1.0, 0.5)) == [
        "BM_Symm<double>/r/l/32/64/1/0.5/usm"]


def test_counters_worked_examples():
    # gemm 42^3, beta 0: 2*42^3 + 42^2 flops; A, B read and C written once
    c = bm.counters("gemm", "float", ("n", "n", 42, 42, 42, 1.0, 0.0))
    assert c["n_fl_ops"] == 149940 and c["bytes_processed"] == 21168 and c["batch_size"] == 1
    # beta != 0 adds 2 m n flops and a read of C; batch multiplies everything
    c = bm.counters("gemm_batched_strided", "double", ("n", "n", 2, 3, 4, 1.0, 0.5, 5, 1, 1, 2))
    assert c["n_fl_ops"] == (48 + 8 + 16) * 5 and c["bytes_processed"] == (6 + 12 + 8 + 8) * 5 * 8
    assert c["stride_c_mul"] == 2
    # gemv 't' 32 x 128, beta 0: x has m = 32 elements, y has n = 128
    c = bm.counters("gemv", "float", ("t", 32, 128, 1.0, 0.0))
    assert c["n_fl_ops"] == 2 * 32 * 128 + 32 and c["bytes_processed"] == (32 * 128 + 32 + 128) * 4
    assert bm.counters("dot", "float", (1000,)) == {"size": 1000, "n_fl_ops": 2000, "bytes_processed": 8004}
    assert bm.counters("nrm2", "float", (1000,))["n_fl_ops"] == 2000
    assert bm.counters("iamax", "double", (1000,)) == {"size": 1000, "n_fl_ops": 1000, "bytes_processed": 8008}
    assert bm.counters("reduction", "float", (513, 8195))["bytes_processed"] == 513 * 8195 * 4
    # trsm m = 4, n = 3: side 'l' solves with the 4 x 4 triangle (10 elements),
    # 2 * 10 * 3 + 12 flops; side 'r' with the 3 x 3 one (6 elements), 2 * 6 * 4 + 12
    c = bm.counters("trsm", "float", ("l", "u", "n", "n", 4, 3, 1.0))
    assert c["n_fl_ops"] == 72 and c["bytes_processed"] == (10 + 12 + 12) * 4
    c = bm.counters("trsm", "double", ("r", "u", "n", "n", 4, 3, 1.0))
    assert c["n_fl_ops"] == 60 and c["bytes_processed"] == (12 + 6 + 12) * 8
    # symm m = 4, n = 3: side 'l', beta 1: 2 * 16 * 3 + 2 * 12 flops, B and C read, C written, 10 of A;
    # side 'r', beta 0: 2 * 3^3 flops (as the reference counts), 6 of A
    c = bm.counters("symm", "float", ("l", "u", 4, 3, 1.0, 1.0))
    assert c["n_fl_ops"] == 120 and c["bytes_processed"] == (24 + 12 + 10) * 4
    c = bm.counters("symm", "float", ("r", "u", 4, 3, 1.0, 0.0))
    assert c["n_fl_ops"] == 54 and c["bytes_processed"] == (12 + 12 + 6) * 4


@pytest.mark.gpu
def test_gpu_every_operator_runs(tmp_path, torch_cuda):
    files = {"gemm": "n,n,64,32,48,1,0\nt,t,33,17,65,1.5,0.5\n", "gemm_batched": "n,t,16,16,16,1,0,4,strided\n",
             "gemm_batched_strided": "n,n,64,64,64,1,0,32,1,1,1\n", "gemv": "n,100,200,1,0\nt,100,200,2,1\n",
             "dot": "4096\n", "asum": "4096\n", "nrm2": "4096\n", "iamax": "4096\n", "iamin": "4096\n",
             "reduction": "64,32\n", "trsm": "l,l,n,n,100,70,1\nr,u,t,u,50,130,2\n",
             "symm": "l,u,100,70,1,0.5\nr,l,50,130,2,0\n"}
    for op, text in files.items():
        out = tmp_path / f"{op}.json"
        assert bm.main([op, f"--csv-param={_csv(tmp_path, text, op + '.csv')}", f"--benchmark_out={out}",
                        "--benchmark_out_format=json", "--benchmark_min_time=0.01"]) == 0
        d = json.loads(out.read_text())
        params = bm.parse_csv(str(tmp_path / f"{op}.csv"), op)
        names = [n for p in params for n in bm.benchmark_names(op, "float", p)]
        assert [b["name"] for b in d["benchmarks"]] == names
        for b in d["benchmarks"]:
            assert b["iterations"] >= 1 and 0 < b["best_event_time"] <= b["avg_event_time"]
            assert b["best_overall_time"] <= b["avg_overall_time"] and b["time_unit"] == "ns"
            assert b["n_fl_ops"] > 0 and b["bytes_processed"] > 0
