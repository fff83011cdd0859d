"""GPU parity of the individual kernels behind the C ABI against the CPU oracle and closed forms.

Tolerances: SpMV and reductions change the summation order, so they are compared against the exact value
of what was multiplied with the textbook forward-error bound gamma_k * sum|a||x| (k = row length); with
integer data every partial sum is exact and the results must be bit-identical.  The prox steps evaluate
the reference's formulas without FMA contraction and must match the oracle bit for bit."""
import math

import numpy as np
import pytest

from conftest import instance_path

pytestmark = pytest.mark.gpu

F64, F32, BF16 = 0, 1, 2
PAIRS = [(F64, F64), (F32, F64), (F32, F32), (BF16, F32), (BF16, F64)]
LANES = [0, 1, 2, 4, 8, 16, 32, 64]  # 0 = LDS-staged row blocks


def ragged_csr(rng, m, n, integer=False):
    lens = rng.integers(0, 40, m)
    lens[rng.integers(0, m, 5)] = 0          # empty rows
    lens[0] = min(n, 3000)                    # one long row
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    if integer:
        val = rng.integers(-8, 9, ptr[-1]).astype(np.float64)
        val[val == 0] = 1.0
    else:
        val = rng.standard_normal(ptr[-1]) * 10.0 ** rng.integers(-3, 3, ptr[-1])
    return ptr, idx, val


def rounded(vals, mdt, oracle):
    if mdt == F64:
        return vals
    if mdt == F32:
        return vals.astype(np.float32).astype(np.float64)
    return oracle.round_bf16(vals)


def _exact_and_bound(ptr, idx, val, x, unit):
    m = len(ptr) - 1
    rows = np.repeat(np.arange(m), np.diff(ptr))
    prod = val * x[idx]
    exact = np.array([math.fsum(prod[ptr[i]:ptr[i + 1]]) for i in range(m)])
    absum = np.bincount(rows, weights=np.abs(prod), minlength=m)
    k = np.maximum(np.diff(ptr), 1)
    return exact, (k * unit / (1 - k * unit)) * absum


@pytest.mark.parametrize("mdt,vdt", PAIRS)
@pytest.mark.parametrize("lanes", LANES)
def test_spmv_error_bound_vs_oracle(hipctx, oracle, mdt, vdt, lanes):
    from ganit_v2_amd import hip
    rng = np.random.default_rng(lanes * 10 + mdt * 3 + vdt)
    m, n = 700, 4000
    ptr, idx, val = ragged_csr(rng, m, n)
    x = rng.standard_normal(n)
    A = hip.DeviceCsr(hipctx, m, n, ptr, idx, val)
    A.set_lanes(lanes)
    xd, yd = hipctx.vector(x, vdt), hipctx.empty(m, vdt)
    A.spmv(xd, yd, mdt)
    y = yd.numpy()
    xr = x if vdt == F64 else x.astype(np.float32).astype(np.float64)
    exact, bound = _exact_and_bound(ptr, idx, rounded(val, mdt, oracle), xr, 2.0 ** (-53 if vdt == F64 else -24))
    assert np.all(np.abs(y - exact) <= bound + 1e-300)
    yo = oracle.spmv(mdt, vdt, m, n, ptr, idx, val, x)
    assert np.all(np.abs(y - yo) <= 2 * bound + 1e-300)
    A.spmv(xd, yd, mdt)  # deterministic
    assert np.array_equal(yd.numpy(), y)


@pytest.mark.parametrize("lanes", LANES)
def test_spmv_integer_data_is_exact(hipctx, lanes):
    from ganit_v2_amd import hip
    rng = np.random.default_rng(100 + lanes)
    m, n = 500, 3000
    ptr, idx, val = ragged_csr(rng, m, n, integer=True)
    x = rng.integers(-50, 51, n).astype(np.float64)
    A = hip.DeviceCsr(hipctx, m, n, ptr, idx, val)
    A.set_lanes(lanes)
    for vdt in (F64, F32):
        xd, yd = hipctx.vector(x, vdt), hipctx.empty(m, vdt)
        A.spmv(xd, yd, F64 if vdt == F64 else F32)
        rows = np.repeat(np.arange(m), np.diff(ptr))
        want = np.bincount(rows, weights=val * x[idx], minlength=m)
        assert np.array_equal(yd.numpy(), want)


@pytest.mark.parametrize("mdt,vdt", PAIRS)
def test_rowblock_spmv_bitwise_vs_oracle(hipctx, oracle, mdt, vdt):
    """Row blocks stage rounded products in LDS and add each row up in row order: the row sums are the CPU
    loop's bits exactly (rows up to 256 entries; longer rows take the block-reduction path)."""
    from ganit_v2_amd import hip
    rng = np.random.default_rng(20 + 3 * mdt + vdt)
    m, n = 3000, 5000
    lens = rng.integers(0, 257, m)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    val = rng.standard_normal(ptr[-1])
    x = rng.standard_normal(n)
    A = hip.DeviceCsr(hipctx, m, n, ptr, idx, val)
    A.set_lanes(0)
    xd, yd = hipctx.vector(x, vdt), hipctx.empty(m, vdt)
    A.spmv(xd, yd, mdt)
    assert np.array_equal(yd.numpy(), oracle.spmv(mdt, vdt, m, n, ptr, idx, val, x))


def _windowed_blocks(ptr, idx, max_rows=256, max_nnz=2048, klong=256, kwin=2048):
    """Staged row blocks of the default partition whose columns qualify for an LDS window (span <= kwin and
    span <= 2 * block nonzeros): (qualifying, staged)."""
    m, r, hit, staged = len(ptr) - 1, 0, 0, 0
    while r < m:
        r0, nz = r, 0
        if ptr[r + 1] - ptr[r] > klong:
            r += 1
            continue
        while r < m and r - r0 < max_rows and ptr[r + 1] - ptr[r] <= klong and nz + ptr[r + 1] - ptr[r] <= max_nnz:
            nz += ptr[r + 1] - ptr[r]
            r += 1
        if nz:
            cols = idx[ptr[r0]:ptr[r]]
            span = int(cols.max() - cols.min()) + 1
            staged += 1
            hit += span <= kwin and span <= 2 * nz
    return hit, staged


@pytest.mark.parametrize("window", ["plain", "swizzled", "off"])
@pytest.mark.parametrize("mdt,vdt", PAIRS)
def test_windowed_rowblock_spmv_bitwise(hipctx, oracle, monkeypatch, mdt, vdt, window):
    """Banded rows: most blocks gather from a copy of their window of x in LDS; with windows on (in either
    layout) and with them switched off, the row sums are the oracle's bits exactly."""
    from ganit_v2_amd import hip
    if window == "off":
        monkeypatch.setenv("GANIT_HIP_WINDOW", "0")
    else:
        monkeypatch.setenv("GANIT_HIP_WINDOW_SWIZZLE", "1" if window == "swizzled" else "0")
    rng = np.random.default_rng(40 + 3 * mdt + vdt)
    m, n = 4000, 9000
    lens = rng.integers(0, 120, m)
    lens[rng.integers(0, m, 7)] = 0
    starts = (np.arange(m) * 2 + rng.integers(0, 50, m)).clip(0, n - 300)
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.sort(s0 + rng.choice(300, k, replace=False)) for s0, k in zip(starts, lens)])
    idx = idx.astype(np.int32)
    hit, staged = _windowed_blocks(ptr, idx)
    assert hit >= staged // 2 and hit > 20, (hit, staged)
    val = rng.standard_normal(ptr[-1])
    x = rng.standard_normal(n)
    A = hip.DeviceCsr(hipctx, m, n, ptr, idx, val)
    A.set_lanes(0)
    xd, yd = hipctx.vector(x, vdt), hipctx.empty(m, vdt)
    A.spmv(xd, yd, mdt)
    assert np.array_equal(yd.numpy(), oracle.spmv(mdt, vdt, m, n, ptr, idx, val, x))


@pytest.mark.parametrize("mdt,vdt", [(F64, F64), (F32, F32), (BF16, F32)])
def test_compact_indices_match_plain(hipctx, oracle, monkeypatch, mdt, vdt):
    """16-bit column offsets (per staged block, long row and split-row chunk, each falling back to 32-bit
    indices where its span does not fit) give the same bits as the plain 32-bit indices, on one matrix with
    banded short rows, long rows (compact and gathered through an LDS window, and too wide) and split rows
    (compact, too-wide and consecutive-column chunks, the last reading no index at all)."""
    from ganit_v2_amd import hip
    rng = np.random.default_rng(60 + mdt)
    n = 300000
    rows = []
    for i in range(1500):  # short rows, banded: staged blocks, most of them windowed
        k = int(rng.integers(0, 60))
        s0 = min(i * 150, n - 1000)
        rows.append(np.sort(s0 + rng.choice(1000, k, replace=False)))
    rows.append(np.sort(rng.choice(2000, 1000, replace=False)))         # long row, compact and windowed
    rows.append(np.sort(rng.choice(n, 1000, replace=False)))            # long row, span > 65535
    rows.append(np.sort(rng.choice(60000, 20000, replace=False)))       # split row, compact chunks
    rows.append(np.sort(rng.choice(n, 9000, replace=False)))            # split row: two too-wide chunks, one compact
    rows.append(np.sort(100000 + rng.choice(200000, 30000, replace=False)))  # split row, compact chunks
    rows.append(np.arange(7000, 27000))                                 # split row, consecutive columns: no index
    rows.append(np.concatenate([np.arange(1000, 5096),                  # split row: a consecutive chunk, then
                                5096 + np.sort(rng.choice(200000, 6000, replace=False))]))  # compact ones
    lens = np.array([len(r) for r in rows])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate(rows).astype(np.int32)
    val = rng.standard_normal(ptr[-1])
    x = rng.standard_normal(n)
    ys = []
    for compact in ("1", "0"):
        monkeypatch.setenv("GANIT_HIP_COMPACT_IDX", compact)
        A = hip.DeviceCsr(hipctx, len(lens), n, ptr, idx, val)
        A.set_lanes(0)
        xd, yd = hipctx.vector(x, vdt), hipctx.empty(len(lens), vdt)
        A.spmv(xd, yd, mdt)
        ys.append(yd.numpy())
    assert np.array_equal(ys[0], ys[1])
    xr = x if vdt == F64 else x.astype(np.float32).astype(np.float64)
    exact, bound = _exact_and_bound(ptr, idx, rounded(val, mdt, oracle), xr, 2.0 ** (-53 if vdt == F64 else -24))
    assert np.all(np.abs(ys[0] - exact) <= bound + 1e-300)


@pytest.mark.parametrize("mdt,vdt", [(F64, F64), (F32, F32), (BF16, F32)])
def test_split_rows_spmv(hipctx, oracle, mdt, vdt):
    """Rows longer than 8192 entries are cut into 4096-entry chunks on separate workgroups and combined in
    chunk order by the last one: within the SpMV error bound, and deterministic."""
    from ganit_v2_amd import hip
    rng = np.random.default_rng(30 + mdt)
    n = 120000
    lens = np.array([100001, 3, 8193, 0, 50000, 8192, 17, 40000, 1])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    val = rng.standard_normal(ptr[-1])
    x = rng.standard_normal(n)
    A = hip.DeviceCsr(hipctx, len(lens), n, ptr, idx, val)
    A.set_lanes(0)
    xd, yd = hipctx.vector(x, vdt), hipctx.empty(len(lens), vdt)
    A.spmv(xd, yd, mdt)
    y = yd.numpy()
    xr = x if vdt == F64 else x.astype(np.float32).astype(np.float64)
    exact, bound = _exact_and_bound(ptr, idx, rounded(val, mdt, oracle), xr, 2.0 ** (-53 if vdt == F64 else -24))
    assert np.all(np.abs(y - exact) <= bound + 1e-300)
    for _ in range(3):
        A.spmv(xd, yd, mdt)
        assert np.array_equal(yd.numpy(), y)


def test_autotune_picks_a_valid_variant(hipctx):
    from ganit_v2_amd import hip
    import ganit_v2_amd as g
    s = g.Problem.generate("blending", crudes=20, products=6, specs=5, periods=30, seed=0).scaled()
    for M in (s["A"], s["AT"]):
        d = hip.DeviceCsr(hipctx, M.rows, M.cols, M.ptr, M.idx, M.val)
        lanes = d.autotune(F32, F32, 3)
        assert lanes in LANES
        x = np.random.default_rng(0).standard_normal(M.cols)
        xd, yd = hipctx.vector(x), hipctx.empty(M.rows)
        d.spmv(xd, yd, F64)
        assert np.allclose(yd.numpy(), M.matvec(x), rtol=1e-12, atol=1e-12)


def test_spmv_plan_from_row_histogram(hipctx):
    from ganit_v2_amd import hip
    import ganit_v2_amd as g
    s = g.Problem.generate("random_lp", seed=0).scaled()
    A = hip.DeviceCsr(hipctx, s["m"], s["n"], s["A"].ptr, s["A"].idx, s["A"].val)
    AT = hip.DeviceCsr(hipctx, s["n"], s["m"], s["AT"].ptr, s["AT"].idx, s["AT"].val)
    la, ha = A.plan()
    lt, ht = AT.plan()
    assert ha.sum() == s["m"] and ht.sum() == s["n"]
    assert lt == 16 and ht[4] == s["n"]          # every A^T row has exactly 10 entries -> bin (8,16]
    assert la == 32                               # rows of A average 25 entries
    blend = g.Problem.generate("blending", crudes=20, products=6, specs=5, periods=30, seed=0).scaled()
    BT = hip.DeviceCsr(hipctx, blend["n"], blend["m"], blend["AT"].ptr, blend["AT"].idx, blend["AT"].val)
    assert BT.plan()[0] == 0                      # short columns (<16 entries) start on the row blocks


@pytest.mark.parametrize("vdt", [F64, F32])
@pytest.mark.parametrize("hasq", [False, True])
def test_primal_step_bitwise(hipctx, oracle, vdt, hasq):
    from ganit_v2_amd import hip
    rng = np.random.default_rng(7)
    n = 10000
    x, aty, qx, c, s0 = (rng.standard_normal(n) for _ in range(5))
    l = np.where(rng.random(n) < 0.2, -np.inf, -rng.random(n))
    u = np.where(rng.random(n) < 0.2, np.inf, rng.random(n))
    cast = (lambda a: a) if vdt == F64 else (lambda a: a.astype(np.float32).astype(np.float64))
    x, aty, qx, c, s0, l, u = map(cast, (x, aty, qx, c, s0, l, u))
    tau = 0.7310585786300049
    d = {k: hipctx.vector(v, vdt) for k, v in dict(x=x, aty=aty, qx=qx, c=c, l=l, u=u, s=s0).items()}
    xn, xb = hipctx.empty(n, vdt), hipctx.empty(n, vdt)
    hip.check(hip.lib().ganit_hip_primal_step(hipctx.h, vdt, n, d["x"].ptr, d["aty"].ptr,
                                              d["qx"].ptr if hasq else None, d["c"].ptr, d["l"].ptr, d["u"].ptr,
                                              tau, xn.ptr, xb.ptr, d["s"].ptr))
    exn, exb, es = oracle.primal_step(vdt, x, aty, qx if hasq else None, c, l, u, tau, s0)
    assert np.array_equal(xn.numpy(), exn)
    assert np.array_equal(xb.numpy(), exb)
    assert np.array_equal(d["s"].numpy(), es)


@pytest.mark.parametrize("vdt", [F64, F32])
def test_dual_step_bitwise(hipctx, oracle, vdt):
    from ganit_v2_amd import hip
    rng = np.random.default_rng(8)
    m = 10000
    y, ax, b, s0 = (rng.standard_normal(m) for _ in range(4))
    kind = rng.integers(0, 4, m)
    lo = np.where((kind == 0) | (kind == 2), b, -np.inf)
    hi = np.where((kind == 0) | (kind == 1), b, np.inf)
    cast = (lambda a: a) if vdt == F64 else (lambda a: a.astype(np.float32).astype(np.float64))
    y, ax, lo, hi, s0 = map(cast, (y, ax, lo, hi, s0))
    sigma = 1.2345678901234567
    d = {k: hipctx.vector(v, vdt) for k, v in dict(y=y, ax=ax, lo=lo, hi=hi, s=s0).items()}
    yn = hipctx.empty(m, vdt)
    hip.check(hip.lib().ganit_hip_dual_step(hipctx.h, vdt, m, d["y"].ptr, d["ax"].ptr, d["lo"].ptr, d["hi"].ptr,
                                            sigma, yn.ptr, d["s"].ptr))
    eyn, es = oracle.dual_step(vdt, y, ax, lo, hi, sigma, s0)
    assert np.array_equal(yn.numpy(), eyn)
    assert np.array_equal(d["s"].numpy(), es)


def _fsum_bound(terms):
    terms = np.asarray(terms, dtype=np.float64)
    return math.fsum(terms), len(terms) * 2.0 ** -52 * math.fsum(np.abs(terms)) + 1e-300


@pytest.mark.parametrize("vdt", [F64, F32])
def test_reductions_vs_fsum(hipctx, vdt):
    import ctypes
    from ganit_v2_amd import hip
    rng = np.random.default_rng(9)
    m = n = 200_003
    cast = (lambda a: a) if vdt == F64 else (lambda a: a.astype(np.float32).astype(np.float64))
    ax, y, aty, qx, c, a, b = (cast(rng.standard_normal(n)) for _ in range(7))
    lo = cast(np.where(rng.random(m) < 0.3, -np.inf, -rng.random(m)))
    hi = cast(np.where(rng.random(m) < 0.3, np.inf, rng.random(m)))
    l = cast(np.where(rng.random(n) < 0.3, -np.inf, -rng.random(n)))
    u = cast(np.where(rng.random(n) < 0.3, np.inf, rng.random(n)))
    rr = cast(0.5 + rng.random(m))
    cr = cast(0.5 + rng.random(n))
    D = {k: hipctx.vector(v, vdt) for k, v in dict(ax=ax, y=y, aty=aty, qx=qx, c=c, a=a, b=b, lo=lo, hi=hi, l=l,
                                                   u=u, rr=rr, cr=cr).items()}
    L, h = hip.lib(), hipctx.h
    o0, o1 = ctypes.c_double(), ctypes.c_double()
    hip.check(L.ganit_hip_primal_res2(h, vdt, m, D["ax"].ptr, D["lo"].ptr, D["hi"].ptr, D["rr"].ptr,
                                      ctypes.byref(o0)))
    want, tol = _fsum_bound(((ax - np.clip(ax, lo, hi)) / rr) ** 2)
    assert abs(o0.value - want) <= tol
    for with_c in (1, 0):
        hip.check(L.ganit_hip_dual_res(h, vdt, n, D["aty"].ptr, D["qx"].ptr, D["c"].ptr, D["l"].ptr, D["u"].ptr,
                                       D["cr"].ptr, with_c, ctypes.byref(o0), ctypes.byref(o1)))
        z = with_c * c - aty + qx
        r = np.where(z > 0, np.where(np.isfinite(l), 0.0, z), np.where(z < 0, np.where(np.isfinite(u), 0.0, z), 0.0))
        ob = np.where((z > 0) & np.isfinite(l), z * np.where(np.isfinite(l), l, 0),
                      np.where((z < 0) & np.isfinite(u), z * np.where(np.isfinite(u), u, 0), 0.0))
        w0, t0 = _fsum_bound((r / cr) ** 2)
        w1, t1 = _fsum_bound(ob)
        assert abs(o0.value - w0) <= t0 and abs(o1.value - w1) <= t1
    hip.check(L.ganit_hip_row_dual_obj(h, vdt, m, D["y"].ptr, D["lo"].ptr, D["hi"].ptr, ctypes.byref(o0)))
    terms = np.where((y > 0) & np.isfinite(lo), y * np.where(np.isfinite(lo), lo, 0),
                     np.where((y < 0) & np.isfinite(hi), y * np.where(np.isfinite(hi), hi, 0), 0.0))
    w, t = _fsum_bound(terms)
    assert abs(o0.value - w) <= t
    hip.check(L.ganit_hip_dot(h, vdt, n, D["a"].ptr, D["b"].ptr, ctypes.byref(o0)))
    w, t = _fsum_bound(a * b)
    assert abs(o0.value - w) <= t
    hip.check(L.ganit_hip_diff2(h, vdt, n, D["a"].ptr, D["b"].ptr, ctypes.byref(o0)))
    w, t = _fsum_bound((a - b) ** 2)
    assert abs(o0.value - w) <= t
    # integer data: exact
    ia = hipctx.vector(np.arange(n, dtype=np.float64) % 97 - 48, vdt)
    hip.check(L.ganit_hip_dot(h, vdt, n, ia.ptr, ia.ptr, ctypes.byref(o0)))
    assert o0.value == float(((np.arange(n) % 97 - 48) ** 2).sum())
    # empty input
    hip.check(L.ganit_hip_dot(h, vdt, 0, ia.ptr, ia.ptr, ctypes.byref(o0)))
    assert o0.value == 0.0


def _check_terms(s, x, y, xa, ya, xr, yr, oracle_like=True):
    """The 22 sums of ganit_hip_pdhg_check computed in numpy (math.fsum)."""
    A, AT, Q = s["A"], s["AT"], s["Q"]
    c, l, u, lo, hi, Rr, Cr = (s[k] for k in ("c", "l", "u", "lo", "hi", "Rres", "Cres"))

    def dual(aty, qx, cw):
        z = cw * c - aty + qx
        r = np.where(z > 0, np.where(np.isfinite(l), 0.0, z), np.where(z < 0, np.where(np.isfinite(u), 0.0, z), 0.0))
        ob = np.where((z > 0) & np.isfinite(l), z * np.nan_to_num(l, posinf=0, neginf=0),
                      np.where((z < 0) & np.isfinite(u), z * np.nan_to_num(u, posinf=0, neginf=0), 0.0))
        return (r / Cr) ** 2, ob

    def rowobj(v):
        return np.where((v > 0) & np.isfinite(lo), v * np.nan_to_num(lo, posinf=0, neginf=0),
                        np.where((v < 0) & np.isfinite(hi), v * np.nan_to_num(hi, posinf=0, neginf=0), 0.0))

    def pres(ax):
        return ((ax - np.clip(ax, lo, hi)) / Rr) ** 2

    qxc = Q.matvec(x) if Q.nnz else 0 * x
    qxa = Q.matvec(xa) if Q.nnz else 0 * x
    t = {}
    t[0], t[1] = dual(AT.matvec(y), qxc, 1.0)
    t[2], t[3] = c * x, x * qxc
    t[4], t[5] = dual(AT.matvec(ya), qxa, 1.0)
    t[6], t[7] = c * xa, xa * qxa
    t[8], t[9] = dual(AT.matvec(y - yr), 0 * x, 0.0)
    t[10], t[11] = (x - xr) ** 2, (xa - xr) ** 2
    t[12], t[13] = pres(A.matvec(x)), rowobj(y)
    t[14], t[15] = pres(A.matvec(xa)), rowobj(ya)
    t[16] = rowobj(y - yr)
    t[17], t[18] = (y - yr) ** 2, (ya - yr) ** 2
    d = x - xr
    t[19] = (_cone_excess(A.matvec(d), lo, hi) / Rr) ** 2
    t[20], t[21] = c * d, (_cone_excess(d, l, u) * Cr) ** 2
    return t


def _cone_excess(d, lo, hi):
    """d's distance, entry by entry, to the recession cone of [lo, hi]: {0}, [0, inf), (-inf, 0] or R."""
    fl, fh = np.isfinite(lo), np.isfinite(hi)
    return np.where(fl & fh, d, np.where(fl, np.minimum(d, 0.0), np.where(fh, np.maximum(d, 0.0), 0.0)))


@pytest.mark.parametrize("name", ["afiro", "25fv47", "port_small", "random_lp"])
def test_pdhg_check_sums_fp64(hipctx, name):
    """All 22 sums of one fused check (current, average, both rays) against numpy on the same FP64 data."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    p = g.Problem.generate("random_lp", seed=1) if name == "random_lp" else g.Problem.read_mps(instance_path(name))
    s = p.scaled()
    pd = hip.DevicePdhg(hipctx, s, mdt_mask=1, vdt_mask=1)
    rng = np.random.default_rng(5)
    m, n = s["m"], s["n"]
    x, xr = rng.standard_normal(n), rng.standard_normal(n)
    y, yr = rng.standard_normal(m), rng.standard_normal(m)
    pd.load(F64, x, y, xr, yr)
    # build running sums with one block of iterations, then read the average back through the check
    pd.iterate(F64, F64, 3, 0.1, 0.2)
    out = pd.check(F64, F64, hip.CHECK_AVG | hip.CHECK_RAY | hip.CHECK_DRAY, 1.0 / 3)
    xc, yc, xr2, yr2 = pd.store(F64)
    # the average the kernel formed: re-derive it by adopting it
    pd.adopt_average(F64)
    xa, ya, _, _ = pd.store(F64)
    assert np.array_equal(xr2, xr) and np.array_equal(yr2, yr)
    terms = _check_terms(s, xc, yc, xa, ya, xr, yr)
    k_row = max(np.diff(s["A"].ptr).max(), np.diff(s["AT"].ptr).max(), 1)
    for k in range(22):
        want = math.fsum(terms[k])
        # SpMV-derived terms carry the row-sum rounding; bound generously by k_row * eps relative
        scale = math.fsum(np.abs(terms[k])) + 1e-300
        assert abs(out[k] - want) <= (k_row + len(terms[k])) * 4e-16 * scale * 50 + 1e-300, k


def test_pdhg_iteration_matches_oracle_steps(hipctx, oracle):
    """One fused iteration = spmv_AT -> primal_step -> spmv_A -> dual_step of the oracle (FP64), within the
    SpMV rounding bound; the prox arithmetic itself is bit-exact (tested above)."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    s = g.Problem.read_mps(instance_path("25fv47")).scaled()
    m, n = s["m"], s["n"]
    pd = hip.DevicePdhg(hipctx, s, 1, 1)
    rng = np.random.default_rng(6)
    x0 = np.clip(rng.standard_normal(n), s["l"], s["u"])
    y0 = rng.standard_normal(m)
    pd.load(F64, x0, y0)
    tau, sigma = 0.3, 0.4
    pd.iterate(F64, F64, 1, tau, sigma)
    x1, y1, _, _ = pd.store(F64)
    AT, A = s["AT"], s["A"]
    aty = oracle.spmv(0, 0, n, m, AT.ptr, AT.idx, AT.val, y0)
    xn, xbar, _ = oracle.primal_step(0, x0, aty, None, s["c"], s["l"], s["u"], tau, np.zeros(n))
    axb = oracle.spmv(0, 0, m, n, A.ptr, A.idx, A.val, xbar)
    yn, _ = oracle.dual_step(0, y0, axb, s["lo"], s["hi"], sigma, np.zeros(m))
    kmax = max(np.diff(A.ptr).max(), np.diff(AT.ptr).max())
    scale_x = tau * (np.abs(AT.val) @ np.ones(1) if False else 1.0)
    assert np.max(np.abs(x1 - xn)) <= 1e-12 * (1 + np.max(np.abs(xn)))
    assert np.max(np.abs(y1 - yn)) <= 1e-11 * (1 + np.max(np.abs(yn)))


@pytest.mark.parametrize("swizzle", ["auto", "0", "1"])
@pytest.mark.parametrize("kind", ["random_lp", "blending"])
def test_rowblock_iteration_is_bit_exact(hipctx, oracle, monkeypatch, kind, swizzle):
    """With A and A^T on row blocks, one fused iteration reproduces the oracle's spmv_AT -> primal_step ->
    spmv_A -> dual_step bit for bit (FP64 and FP32); the blending LP's blocks gather through LDS windows, in
    the layout chosen at upload and in each layout forced."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    if swizzle != "auto":
        monkeypatch.setenv("GANIT_HIP_WINDOW_SWIZZLE", swizzle)
    if kind == "random_lp":
        s = g.Problem.generate("random_lp", seed=4).scaled()
    else:
        s = g.Problem.generate("blending", crudes=20, products=6, specs=5, periods=30, seed=0).scaled()
        hit, staged = _windowed_blocks(s["A"].ptr, s["A"].idx)
        assert hit == staged, (hit, staged)
    assert max(np.diff(s["A"].ptr).max(), np.diff(s["AT"].ptr).max()) <= 256
    m, n = s["m"], s["n"]
    pd = hip.DevicePdhg(hipctx, s, mdt_mask=3, vdt_mask=3)
    pd.set_lanes(0, 0)
    pd.set_lanes(1, 0)
    if swizzle != "auto":
        assert pd.window_layout(0) == pd.window_layout(1) == int(swizzle)
    rng = np.random.default_rng(6)
    for vdt, mdt in ((F64, F64), (F32, F32)):
        cast = (lambda a: a) if vdt == F64 else (lambda a: a.astype(np.float32).astype(np.float64))
        x0 = cast(np.clip(rng.standard_normal(n), s["l"], s["u"]))
        y0 = cast(rng.standard_normal(m))
        pd.load(vdt, x0, y0)
        tau, sigma = 0.3, 0.4
        pd.iterate(mdt, vdt, 1, tau, sigma)
        x1, y1, _, _ = pd.store(vdt)
        AT, A = s["AT"], s["A"]
        aty = oracle.spmv(mdt, vdt, n, m, AT.ptr, AT.idx, AT.val, y0)
        xn, xbar, _ = oracle.primal_step(vdt, x0, aty, None, s["c"], s["l"], s["u"], tau, np.zeros(n))
        axb = oracle.spmv(mdt, vdt, m, n, A.ptr, A.idx, A.val, xbar)
        yn, _ = oracle.dual_step(vdt, y0, axb, s["lo"], s["hi"], sigma, np.zeros(m))
        assert np.array_equal(x1, xn) and np.array_equal(y1, yn)


def test_window_layout_follows_the_bank_pattern(hipctx):
    """The blending LP's quality rows gather columns 16 apart (one per crude): its A stages swizzled windows,
    while A^T's gathers (a column's few rows) rarely meet on one bank and keep the plain layout."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    s = g.Problem.generate("blending", crudes=20, products=16, specs=5, periods=30, seed=0).scaled()
    pd = hip.DevicePdhg(hipctx, s, mdt_mask=3, vdt_mask=3)
    assert pd.window_layout(0) == 1 and pd.window_layout(1) == 0


def test_graph_replay_matches_eager(hipctx, monkeypatch):
    """Iteration blocks replayed from a captured HIP graph give the same bits as eager launches, on every
    replay, with host work and synchronisations in between."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    s = g.Problem.generate("random_lp", seed=2).scaled()
    rng = np.random.default_rng(3)
    x0 = np.clip(rng.standard_normal(s["n"]), s["l"], s["u"])
    y0 = rng.standard_normal(s["m"])
    pg = hip.DevicePdhg(hipctx, s, 1, 1)
    monkeypatch.setenv("GANIT_HIP_NO_GRAPH", "1")
    pe = hip.DevicePdhg(hipctx, s, 1, 1)
    for pd in (pg, pe):  # same variants in both (the autotuner may pick differently per object)
        pd.set_lanes(0, 0)
        pd.set_lanes(1, 8)
        pd.load(F64, x0, y0)
    for rep in range(4):
        for pd in (pg, pe):
            pd.iterate(F64, F64, 64, 0.05 * (rep + 1), 0.07)
        hipctx.sync()
        a, b = pg.store(F64), pe.store(F64)
        for u, v in zip(a, b):
            assert np.array_equal(u, v), rep
        oa = pg.check(F64, F64, hip.CHECK_AVG, 1.0 / (64 * (rep + 1)))
        ob = pe.check(F64, F64, hip.CHECK_AVG, 1.0 / (64 * (rep + 1)))
        assert np.array_equal(oa, ob)


@pytest.mark.parametrize("lanes", [0, 4])
def test_fp64_check_of_fp32_iterate(hipctx, lanes):
    """GANIT_CHECK_FP64 on an FP32 working set equals the FP64 check of the same (widened) iterate."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    s = g.Problem.read_mps(instance_path("25fv47")).scaled()
    pd = hip.DevicePdhg(hipctx, s, mdt_mask=3, vdt_mask=3)
    for which in (0, 1):
        pd.set_lanes(which, lanes)
    rng = np.random.default_rng(11)
    x = np.clip(rng.standard_normal(s["n"]), s["l"], s["u"]).astype(np.float32).astype(np.float64)
    y = rng.standard_normal(s["m"]).astype(np.float32).astype(np.float64)
    pd.load(F32, x, y)
    pd.load(F64, x, y)
    a = pd.check(F64, F32, hip.CHECK_RAY | hip.CHECK_FP64)
    b = pd.check(F64, F64, hip.CHECK_RAY)
    assert np.array_equal(a, b)
    c = pd.check(F32, F32, hip.CHECK_RAY)   # the level's own evaluation differs (FP32 sums)
    assert not np.array_equal(a, c)
    # with the average: after one FP32 iteration the running sums are the iterate itself, so with inv_k = 1 the
    # FP32 average equals the iterate and its FP64 evaluation equals the current point's
    pd.load(F32, x, y)
    pd.iterate(F32, F32, 1, 0.3, 0.4)
    x1, y1, xr, yr = pd.store(F32)
    pd.load(F64, x1, y1, xr, yr)
    a = pd.check(F64, F32, hip.CHECK_AVG | hip.CHECK_RAY | hip.CHECK_DRAY | hip.CHECK_FP64, 1.0)
    b = pd.check(F64, F64, hip.CHECK_RAY | hip.CHECK_DRAY)
    cur = [0, 1, 2, 3, 8, 9, 12, 13, 16, 19, 20, 21]  # entries of the current point and the two rays
    assert np.array_equal(a[cur], b[cur])
    assert np.array_equal(a[4:8], a[0:4]) and np.array_equal(a[14:16], a[12:14])
    assert a[11] == a[10] and a[18] == a[17] and a[10] > 0


def test_handover_and_restart_semantics(hipctx):
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    s = g.Problem.read_mps(instance_path("afiro")).scaled()
    pd = hip.DevicePdhg(hipctx, s, mdt_mask=7, vdt_mask=3)
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(s["n"]), rng.standard_normal(s["m"])
    xr, yr = rng.standard_normal(s["n"]), rng.standard_normal(s["m"])
    pd.load(F64, x, y, xr, yr)
    pd.handover(F64, F32)
    for got, want in zip(pd.store(F32), (x, y, xr, yr)):
        assert np.array_equal(got, want.astype(np.float32).astype(np.float64))
    pd.iterate(BF16, F32, 5, 0.1, 0.1)
    pd.check(BF16, F32, hip.CHECK_AVG, 0.2)
    pd.adopt_average(F32)
    xa, ya, xr1, yr1 = pd.store(F32)
    pd.restart(F32, 0)
    x2, y2, xr2, yr2 = pd.store(F32)
    assert np.array_equal(x2, xa) and np.array_equal(xr2, xa) and np.array_equal(yr2, ya)
    out = pd.check(BF16, F32, hip.CHECK_AVG, 1.0)
    assert out[10] == 0.0 and out[17] == 0.0      # iterate == restart point
    assert out[4] == 0 or np.isfinite(out[4])


@pytest.mark.parametrize("name", ["25fv47", "random_lp"])
def test_pdhg_check_sums_kahan_fp32(hipctx, name):
    """GANIT_CHECK_KAHAN: the same 19 sums from FP32-rounded terms with Kahan-compensated FP32 accumulation,
    against the FP64-accumulated sums of the same FP32 iterate: within a few FP32 roundings of sum |terms|."""
    import ganit_v2_amd as g
    from ganit_v2_amd import hip
    p = g.Problem.generate("random_lp", seed=2) if name == "random_lp" else g.Problem.read_mps(instance_path(name))
    s = p.scaled()
    pd = hip.DevicePdhg(hipctx, s, mdt_mask=2, vdt_mask=2)
    rng = np.random.default_rng(8)
    m, n = s["m"], s["n"]
    pd.load(F32, rng.standard_normal(n), rng.standard_normal(m), rng.standard_normal(n), rng.standard_normal(m))
    pd.iterate(F32, F32, 4, 0.1, 0.2)
    flags = hip.CHECK_AVG | hip.CHECK_RAY
    # the check forms the average from the running sums; both calls below see the same sums and iterate
    f64 = pd.check(F32, F32, flags, 0.25)
    kah = pd.check(F32, F32, flags | hip.CHECK_KAHAN, 0.25)
    xc, yc, xr, yr = pd.store(F32)
    pd.adopt_average(F32)
    xa, ya, _, _ = pd.store(F32)
    terms = _check_terms(s, xc, yc, xa, ya, xr, yr)
    for k in range(19):
        scale = math.fsum(np.abs(terms[k])) + 1e-30
        assert abs(kah[k] - f64[k]) <= 1e-6 * scale, (k, kah[k], f64[k], scale)
    assert not np.array_equal(kah, f64)  # the FP32 path is really taken
