"""Host library (libganit.so): reader, writer, scaling and generators, checked against the oracle's
restatement of the reference and against closed forms.  CPU only."""
import os

import numpy as np
import pytest

import ganit_v2_amd as g
from conftest import instance_path

ALL = ["afiro", "adlittle", "bgetam", "etamacro", "stair", "shell", "25fv47", "greenbea", "port_tiny",
       "port_small", "qblend_small"]


def _same_lp(p: g.Problem, o) -> None:
    a = p.export()
    b = o.export()
    assert np.array_equal(a["A"].ptr, b["aptr"]) and np.array_equal(a["A"].idx, b["aidx"])
    assert np.array_equal(a["A"].val, b["aval"])
    if a["Q"].nnz:
        assert np.array_equal(a["Q"].ptr, b["qptr"]) and np.array_equal(a["Q"].idx, b["qidx"])
        assert np.array_equal(a["Q"].val, b["qval"])
    else:
        assert len(b["qval"]) == 0
    for k in ("c", "l", "u", "lo", "hi"):
        assert np.array_equal(a[k], b[k]), k
    assert a["obj_const"] == b["obj_const"]


@pytest.mark.parametrize("name", ALL)
def test_reader_matches_oracle(oracle, name):
    _same_lp(g.Problem.read_mps(instance_path(name)), oracle.OracleLP.read_mps(instance_path(name)))


@pytest.mark.parametrize("name", ALL)
def test_scaling_bitwise_equal_to_oracle(oracle, name):
    s = g.Problem.read_mps(instance_path(name)).scaled()
    o = oracle.OracleLP.read_mps(instance_path(name)).scaled()
    for a, b in [("A", "a"), ("AT", "t"), ("Q", "q")]:
        assert np.array_equal(s[a].ptr, o[b + "ptr"]) and np.array_equal(s[a].idx, o[b + "idx"])
        assert np.array_equal(s[a].val, o[b + "val"]), a
    for k in ("c", "l", "u", "lo", "hi", "R", "C", "Rres", "Cres"):
        assert np.array_equal(s[k], o[k]), k
    for k in ("bscale", "cscale", "normQ", "obj_const", "bnorm", "cnorm", "normA"):
        assert s[k] == o[k], k


@pytest.mark.parametrize("kind,params", [("blending", dict(periods=24)), ("random_lp", dict(m=3000, n=7000)),
                                         ("portfolio", dict(assets=3000, factors=20))])
def test_scaling_of_generated_problems_bitwise_equal_to_oracle(oracle, kind, params):
    """The parallel column view, A^T and power iteration of the scaling on larger generated problems (uneven row
    lengths, many threads' row ranges): every array and scalar equals the oracle's serial computation."""
    p = g.Problem.generate(kind, seed=3, **params)
    s = p.scaled()
    o = oracle.OracleLP.from_export(p.export()).scaled()
    for a, b in [("A", "a"), ("AT", "t"), ("Q", "q")]:
        assert np.array_equal(s[a].ptr, o[b + "ptr"]) and np.array_equal(s[a].idx, o[b + "idx"]), a
        assert np.array_equal(s[a].val, o[b + "val"]), a
    for k in ("c", "l", "u", "lo", "hi", "R", "C", "Rres", "Cres"):
        assert np.array_equal(s[k], o[k]), k
    for k in ("bscale", "cscale", "normQ", "obj_const", "bnorm", "cnorm", "normA"):
        assert s[k] == o[k], k


@pytest.mark.parametrize("name", ["afiro", "bgetam", "port_tiny", "qblend_small", "etamacro"])
def test_mps_write_read_round_trip(tmp_path, name):
    p = g.Problem.read_mps(instance_path(name))
    f = str(tmp_path / "rt.mps")
    p.write_mps(f)
    q = g.Problem.read_mps(f)
    a, b = p.export(), q.export()
    for k in ("c", "l", "u", "lo", "hi"):
        assert np.array_equal(a[k], b[k]), k
    for M in ("A", "Q"):
        assert np.array_equal(a[M].ptr, b[M].ptr) and np.array_equal(a[M].idx, b[M].idx)
        assert np.array_equal(a[M].val, b[M].val)
    assert a["obj_const"] == b["obj_const"] and a["maximize"] == b["maximize"]


def test_reader_number_formats(tmp_path, oracle):
    """Coefficients in every decimal form the reader's fast path takes or hands to strtod (signs, leading and
    trailing zeros, 1-25 significant digits, exponents up to +-330, 2^53 + 1, subnormals, '1.', '.5', '-0'):
    the values equal Python's correctly rounded float() and the oracle's strtod, bit for bit."""
    rng = np.random.default_rng(12)
    forms = ["1", "1.", ".5", "-.5", "+2.5", "-0", "-0.0", "0.1", "0.30000000000000004", "1e22", "1e23", "1E-05",
             "5e3", "1.0e-5", "123456789012345678", "1234567890123456789", "12345678901234567890",
             "9007199254740992", "9007199254740993", "0.000000000000000000000000001", "4.9e-324", "2.2250738585072014e-308",
             "1.7976931348623157e308", "000123.4500", "-00.00012e+3", "3.14159265358979323846264", "1e-22", "8.5e-23"]
    for _ in range(400):
        digits = "".join(rng.choice(list("0123456789"), int(rng.integers(1, 26))))
        dot = int(rng.integers(0, len(digits) + 1))
        num = digits[:dot] + "." + digits[dot:] if rng.random() < 0.7 else digits
        if rng.random() < 0.5:
            num += ("e" if rng.random() < 0.5 else "E") + str(int(rng.integers(-330, 330)))
        forms.append(("-" if rng.random() < 0.3 else "") + num)
    forms = [f for f in forms if float(f) != 0.0 and np.isfinite(float(f))] + ["-0", "-0.0"]
    n = len(forms)
    lines = ["NAME NUMBERS", "ROWS", " N OBJ", " E R1", "COLUMNS"]
    for j, f in enumerate(forms):
        lines.append(f"    X{j} OBJ {f} R1 {forms[(j + 1) % n]}")
    lines += ["RHS", "    RHS R1 1", "BOUNDS", " UP BND X0 4", "ENDATA", ""]
    path = tmp_path / "numbers.mps"
    path.write_text("\n".join(lines))
    p = g.Problem.read_mps(str(path))
    _same_lp(p, oracle.OracleLP.read_mps(str(path)))
    d = p.export()
    want_c = np.array([float(f) for f in forms]) + 0.0
    assert np.array_equal(d["c"], want_c)
    want_a = np.array([float(forms[(j + 1) % n]) for j in range(n)])
    keep = want_a != 0.0
    assert np.array_equal(d["A"].val, want_a[keep])


def _big_mps(rng, n_rows=2000, n_cols=60000):
    """A multi-megabyte free-MPS text: columns spread over several lines, repeated objective entries, entries in
    an extra N row, duplicate entries, integer MARKER ranges, ranged rows and bounds."""
    lines = ["NAME BIG", "ROWS", " N COST", " N COST2"]
    types = rng.choice(list("ELG"), n_rows)
    lines += [f" {t} R{i}" for i, t in enumerate(types)]
    lines.append("COLUMNS")
    in_int = False
    for j in range(n_cols):
        if j % 997 == 0:
            lines.append(f"    MARKER                 'MARKER'                 '{'INTEND' if in_int else 'INTORG'}'")
            in_int = not in_int
        k = int(rng.integers(1, 7))
        rows = list(rng.integers(0, n_rows, k))
        if rng.random() < 0.05:
            rows.append(rows[0])  # a duplicate entry: summed
        ent = [(f"R{r}", f"{rng.normal():.15g}") for r in rows]
        if rng.random() < 0.7:
            ent.insert(int(rng.integers(0, len(ent) + 1)), ("COST", f"{rng.uniform(1, 10):.12g}"))
        if rng.random() < 0.1:
            ent.append(("COST", f"{rng.normal():.9g}"))  # a second objective entry, added in line order
        if rng.random() < 0.1:
            ent.append(("COST2", "3.5"))  # extra N row: ignored
        for a in range(0, len(ent), 2):
            pair = ent[a:a + 2]
            lines.append(f"    X{j}  " + "  ".join(f"{r} {v}" for r, v in pair))
    lines.append("RHS")
    lines += [f"    RHS R{i} {rng.normal():.12g}" for i in range(0, n_rows, 3)]
    lines.append("RANGES")
    lines += [f"    RNG R{i} {abs(rng.normal()):.6g}" for i in range(1, n_rows, 7)]
    lines.append("BOUNDS")
    lines += [f" UP BND X{j} {rng.uniform(1, 5):.6g}" for j in range(0, n_cols, 11)]
    lines += [f" LO BND X{j} -2" for j in range(5, n_cols, 13)]
    lines.append("ENDATA")
    return lines


def test_reader_large_file_matches_oracle(tmp_path, oracle):
    """A file large enough for the parallel COLUMNS parse (many chunks, columns split across chunk boundaries):
    the same LP as the oracle's sequential restatement, and the sequential loop's first error and line."""
    rng = np.random.default_rng(21)
    lines = _big_mps(rng)
    path = tmp_path / "big.mps"
    path.write_text("\n".join(lines) + "\n")
    assert path.stat().st_size > 4 << 20
    _same_lp(g.Problem.read_mps(str(path)), oracle.OracleLP.read_mps(str(path)))
    # an unknown row and, later, a bad number deep in COLUMNS: the first one in file order is reported
    col0 = lines.index("COLUMNS")
    bad_at = col0 + 1 + (len(lines) - col0) * 2 // 3
    lines.insert(bad_at, "    X999999  NOPE 1.0")
    lines.insert(bad_at + 5000, "    X999998  R1 ABC")  # no numeric prefix (strtod would take "1.0x" as 1.0)
    path.write_text("\n".join(lines) + "\n")
    with pytest.raises(RuntimeError, match=f"unknown row 'NOPE' at line {bad_at + 1}"):
        g.Problem.read_mps(str(path))
    del lines[bad_at]
    path.write_text("\n".join(lines) + "\n")
    with pytest.raises(RuntimeError, match=f"MPS line {bad_at + 5000}: bad number 'ABC'"):
        g.Problem.read_mps(str(path))


def test_reader_edge_cases(tmp_path, oracle):
    """RANGES on E/L/G rows, negative UP on a default lower bound, objective constant, extra N rows,
    MARKER lines, OBJSENSE MAX, duplicate entries and explicit zeros -- reader vs oracle restatement."""
    text = """NAME EDGE
OBJSENSE
    MAX
ROWS
 N COST
 N OTHER
 E R1
 L R2
 G R3
 E R4
COLUMNS
    X1 COST 1 R1 1
    X1 R2 2 OTHER 5
    MARKER 'MARKER' 'INTORG'
    X2 COST -1 R1 0
    X2 R3 1 R2 1
    X2 R2 1
    MARKER 'MARKER' 'INTEND'
    X3 R4 1 R3 -1
RHS
    RHS COST 3.5 R1 1
    RHS R2 4 R3 -2
    RHS R4 2
RANGES
    RNG R1 2 R2 3
    RNG R3 -1 R4 -4
BOUNDS
 UP BND X1 -3
 LO BND X2 -1
 UP BND X2 -0.5
 FR BND X3
ENDATA
"""
    f = tmp_path / "edge.mps"
    f.write_text(text)
    p = g.Problem.read_mps(str(f))
    _same_lp(p, oracle.OracleLP.read_mps(str(f)))
    e = p.export()
    assert e["maximize"] and e["obj_const"] == -3.5
    assert e["l"][0] == -np.inf and e["u"][0] == -3          # negative UP frees the default lower bound
    assert e["l"][1] == -1 and e["u"][1] == -0.5            # explicit LO is kept
    assert e["l"][2] == -np.inf and e["u"][2] == np.inf
    assert list(e["lo"]) == [1, 1, -2, -2] and list(e["hi"]) == [3, 4, -1, 2]
    A = e["A"].to_dense()
    assert A[1, 1] == 2.0                                   # duplicates summed
    assert A[0, 1] == 0.0 and e["A"].nnz == 6               # explicit zero dropped


def test_random_lp_generator_planted_point():
    p = g.Problem.generate("random_lp", seed=3)
    e = p.export()
    x0, y0 = p.planted()
    A = e["A"]
    assert (A.rows, A.cols, A.nnz) == (2000, 5000, 50000)
    cols = np.bincount(A.idx, minlength=5000)
    assert np.all(cols == 10)                               # 10 distinct rows per column
    ax = A.matvec(x0)
    assert np.allclose(ax[:1000], e["lo"][:1000], rtol=0, atol=1e-9)
    assert np.array_equal(e["lo"][:1000], e["hi"][:1000])
    slack = e["hi"][1000:] - ax[1000:]
    assert np.all(slack >= -1e-9) and np.all(slack <= 1 + 1e-9)
    assert np.all(np.isneginf(e["lo"][1000:]))
    assert np.all((x0 >= 0) & (x0 <= 10)) and abs(np.mean(x0 > 0) - 0.7) < 0.03
    assert np.all(y0[1000:] <= 0)
    # c = A^T y0 + s0 with s0 >= 0 only where x0 == 0
    s0 = e["c"] - np.bincount(A.idx, weights=A.val * y0[np.repeat(np.arange(2000), np.diff(A.ptr))], minlength=5000)
    assert np.all(s0 > -1e-9) and np.all(np.abs(s0[x0 > 0]) < 1e-9)
    # weak duality bracket: dual objective of (y0, s0) <= c'x0
    dual = y0[:1000] @ e["lo"][:1000] + y0[1000:] @ e["hi"][1000:] + s0 @ e["l"]
    assert dual <= e["c"] @ x0 + 1e-6


def test_generators_are_deterministic():
    a = g.Problem.generate("random_lp", seed=7, m=300, n=700, m_eq=150).export()
    b = g.Problem.generate("random_lp", seed=7, m=300, n=700, m_eq=150).export()
    c = g.Problem.generate("random_lp", seed=8, m=300, n=700, m_eq=150).export()
    assert np.array_equal(a["A"].val, b["A"].val) and np.array_equal(a["c"], b["c"])
    assert not np.array_equal(a["A"].val, c["A"].val)


def test_random_lp_tightening_and_ragged_columns():
    p = g.Problem.generate("random_lp", seed=11, m=400, n=1500, m_eq=200, nnz_min=5, nnz_max=15,
                           tighten_frac=0.05)
    e = p.export()
    x0, _ = p.planted()
    k = np.bincount(e["A"].idx, minlength=1500)
    assert k.min() >= 5 and k.max() <= 15
    tight = (e["u"] < 10) | (e["l"] > 0)
    assert 0.02 < tight.mean() < 0.09
    assert np.all(e["u"][e["u"] < 10] == np.floor(x0[e["u"] < 10]))
    assert np.all(e["l"][e["l"] > 0] == np.ceil(x0[e["l"] > 0]))


def test_blending_generator_structure():
    p = g.Problem.generate("blending", crudes=6, products=3, specs=5, periods=4, seed=1)
    d = p.dims()
    C, P, K, T = 6, 3, 5, 4
    assert d["n"] == T * (C * P + 2 * C + P)
    assert d["m"] == T * (P + 2 * C + P * K + 1)
    # per x variable: demand, carry-over, capacity and K specs, minus the one exact-zero spec coefficient
    # per (spec, product, period) that the CSR build drops
    # (r: availability + carry-over; ci and pi: two balance rows except in the last period)
    full = T * (C * P * (3 + K) + 2 * C + 2 * C + 2 * P) - C - P
    assert d["nnz"] == full - K * P * T


def test_blending_full_size():
    d = g.Problem.generate("blending", seed=0).dims()
    assert d["m"] == 365 * 217 and d["n"] == 365 * 1096
    assert 2_800_000 <= d["nnz"] <= 2_950_000


def test_portfolio_generator_structure():
    p = g.Problem.generate("portfolio", assets=2000, factors=10, seed=0)
    e = p.export()
    d = p.dims()
    assert (d["m"], d["n"], d["nnz"], d["qnnz"]) == (11, 2010, 2000 * 11 + 10, 2010)
    assert np.all(e["u"][:2000] == 0.01) and np.all(np.isinf(e["l"][2000:]))
    assert np.all(e["Q"].val > 0)


def test_diff_qp_generator_convex():
    p = g.Problem.generate("diff_qp", n=300, m_eq=40, m_le=30, seed=5)
    e = p.export()
    Q = e["Q"].to_dense()
    assert np.array_equal(Q, Q.T)
    off = np.abs(Q).sum(1) - np.abs(np.diag(Q))
    assert np.all(np.diag(Q) > off)                       # strictly diagonally dominant -> positive definite
    x0, _ = p.planted()
    ax = e["A"].matvec(x0)
    assert np.allclose(ax[:40], e["lo"][:40], atol=1e-9)
    assert np.all(e["hi"][40:] - ax[40:] >= -1e-9)


def test_batch_workload_generator_sizes():
    """The batched-LP workload: sizes log-uniform in [200, 4000] x [500, 10000], half the rows equalities,
    column counts 5..15, deterministic per seed, and each instance the random-LP generator at its seed."""
    import math
    ms, ns = [], []
    for s in range(64):
        p = g.Problem.generate("batch_lp", seed=s)
        d = p.dims()
        assert 200 <= d["m"] <= 4000 and 500 <= d["n"] <= 10000
        e = p.export()
        col_counts = np.bincount(e["A"].idx, minlength=d["n"])
        assert col_counts.min() >= 5 and col_counts.max() <= 15
        assert np.sum(np.isneginf(e["lo"])) == d["m"] - d["m"] // 2  # the <= rows
        ms.append(math.log(d["m"]))
        ns.append(math.log(d["n"]))
        q = g.Problem.generate("random_lp", seed=s, m=d["m"], n=d["n"], m_eq=d["m"] // 2, nnz_min=5, nnz_max=15,
                               tighten_frac=0.05)
        assert np.array_equal(q.export()["A"].val, e["A"].val) and np.array_equal(q.export()["c"], e["c"])
    # log-uniform: mean of log within a few standard errors of the interval's midpoint
    assert abs(np.mean(ms) - 0.5 * (math.log(200) + math.log(4000))) < 0.35
    assert abs(np.mean(ns) - 0.5 * (math.log(500) + math.log(10000))) < 0.35
    a = g.Problem.generate("batch_lp", seed=7).export()
    b = g.Problem.generate("batch_lp", seed=7).export()
    assert np.array_equal(a["A"].val, b["A"].val) and np.array_equal(a["u"], b["u"])


def test_blending_generator_writes_the_golden_lp(tmp_path):
    """The blending LP the reference CLI's golden runs were made on (tests/golden/reference_runs.jsonl, meta line
    of blending_s0) is what the generator still produces: same dimensions and the same MPS bytes (sha256)."""
    import hashlib
    import json
    from conftest import GOLDEN
    with open(os.path.join(GOLDEN, "reference_runs.jsonl")) as f:
        meta = next(json.loads(l)["meta"] for l in f if '"blending_s0"' in l and '"meta"' in l)
    p = g.Problem.generate("blending", seed=0, crudes=meta["crudes"], products=meta["products"], specs=meta["specs"],
                           periods=meta["periods"])
    d = p.dims()
    assert (d["m"], d["n"], d["nnz"]) == (meta["m"], meta["n"], meta["nnz"])
    path = tmp_path / "blending.mps"
    p.write_mps(str(path))
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    assert h.hexdigest()[:16] == meta["mps_sha256_16"]


def test_portfolio_generator_writes_the_golden_qp(tmp_path):
    """The portfolio QP the reference CLI's golden runs were made on (meta line of portfolio_s0) is what the
    generator still produces: same dimensions and the same QPS bytes (sha256)."""
    import hashlib
    import json
    from conftest import GOLDEN
    with open(os.path.join(GOLDEN, "reference_runs.jsonl")) as f:
        meta = next(json.loads(l)["meta"] for l in f if '"portfolio_s0"' in l and '"meta"' in l)
    p = g.Problem.generate("portfolio", seed=0, assets=meta["assets"], factors=meta["factors"])
    d = p.dims()
    assert (d["m"], d["n"], d["nnz"], d["qnnz"]) == (meta["m"], meta["n"], meta["nnz"], meta["qnnz"])
    path = tmp_path / "portfolio.mps"
    p.write_mps(str(path))
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    assert h.hexdigest()[:16] == meta["mps_sha256_16"]
