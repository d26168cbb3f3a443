"""Kernel-level parity of the HIP path (through the psh_* C-ABI) with the oracle, on the GPU.

fp64 and fp32 (reference "mixed") results of the reference Backend operations must be
bit-identical to the CPU restatement: row-sequential SpMV in CSR order, the r²HPDHG steps with
the same expression order, CPU-order reductions and KKT sums. The fused chunk must equal the
unfused op sequence bit for bit, and the fused check's sums must equal the separate reductions.
"""
import numpy as np
import pytest

from conftest import netlib

pytestmark = pytest.mark.gpu


def long_row_model():
    """Row 0 touches all 3000 columns and column 0 all 1500 rows (both beyond one LDS tile)."""
    from ps26119_amd import lpgen
    m, n = 1500, 3000
    rows, cols, vals = [], [], []
    for j in range(n):
        rows.append(0), cols.append(j), vals.append(1.0 + (j % 5) * 0.25)
        i = 1 + j % (m - 1)
        rows.append(i), cols.append(j), vals.append(0.5 + (j % 7) * 0.1)
        if j == 0:
            for r in range(1, m):
                if r != i:
                    rows.append(r), cols.append(0), vals.append(0.3)
    cs, ri, v = lpgen._csc_from_triplets(m, n, np.array(rows), np.array(cols), np.array(vals))
    c = -1.0 - (np.arange(n) % 13) * 0.1
    return lpgen.Lp(m=m, n=n, c=c, col_lower=np.zeros(n), col_upper=np.full(n, 10.0), row_lower=np.full(m, -np.inf),
                    row_upper=np.full(m, 50.0), col_start=cs, row_index=ri, value=v, name="long-rows")


def models():
    from ps26119_amd import lpgen
    return {"afiro": netlib("afiro"), "stocfor1": netlib("stocfor1"), "rand300": lpgen.random_lp(300, seed=3),
            "long_rows": long_row_model(), "rand20k": lpgen.random_lp(20000, seed=5)}


MODELS = ["afiro", "stocfor1", "rand300", "long_rows", "rand20k"]
_cache = {}


def model(name):
    if name not in _cache:
        _cache[name] = models()[name]
    return _cache[name]


def context(oracle, lp, prec):
    from ps26119_amd import psh
    ctx = psh.Context(psh.scaled_problem(oracle.scale(lp), lp))
    ctx.set_fp32(prec)
    return ctx


def run_ops(ctx, lp, tau=0.3, sigma=0.4, rho=1.0, iters=20):
    from ps26119_amd import psh
    ctx.upload(psh.X, np.sin(1.0 + np.arange(lp.n)))
    ctx.upload(psh.Y, np.cos(2.0 + np.arange(lp.m)))
    ctx.copy(psh.X0, psh.X)
    ctx.copy(psh.Y0, psh.Y)
    for it in range(iters):
        w = (it + 1.0) / (it + 2.0)
        ctx.spmv_t(psh.Y, psh.ATY)
        ctx.r2h_primal(psh.X, psh.X0, psh.ATY, tau, w, rho, psh.XH, psh.XB)
        ctx.spmv(psh.XB, psh.AX)
        ctx.r2h_dual(psh.Y, psh.Y0, psh.AX, sigma, w, rho, psh.YH, psh.YB)


@pytest.mark.parametrize("name", MODELS)
@pytest.mark.parametrize("prec", [0, 1])
def test_backend_ops_bit_identical(gpu_available, oracle, name, prec):
    from ps26119_amd import psh
    lp = model(name)
    ctx = context(oracle, lp, prec)
    run_ops(ctx, lp)
    ctx.primal_step(psh.X, psh.ATY, 0.15, psh.XB)
    ctx.dual_step(psh.Y, psh.AX, 0.5, psh.YB)
    ctx.axpby(0.3, psh.XH, -1.7, psh.XB)
    got = {k: ctx.download(s) for k, s in [("x", psh.X), ("xh", psh.XH), ("xb", psh.XB), ("y", psh.Y), ("yh", psh.YH),
                                            ("yb", psh.YB)]}
    red = [ctx.dot(psh.X, psh.XH, psh.RED_CPU), np.sqrt(ctx.dot(psh.Y, psh.Y, psh.RED_CPU)),
           ctx.diff_norm2(psh.X, psh.X0, psh.RED_CPU)] + ctx.kkt(psh.XH, psh.YH, psh.RED_CPU)
    exp, ered = oracle.backend_ops(lp, precision=prec, iters=20, tau=0.3, sigma=0.4, rho=1.0)
    for k in got:
        assert np.array_equal(got[k], exp[k]), (k, float(np.max(np.abs(got[k] - exp[k]))))
    assert [float(v).hex() for v in red] == [float(v).hex() for v in ered]
    # tree reductions: deterministic and within rounding of the CPU order
    tree = [ctx.dot(psh.X, psh.XH), np.sqrt(ctx.dot(psh.Y, psh.Y)), ctx.diff_norm2(psh.X, psh.X0)] + ctx.kkt(psh.XH, psh.YH)
    again = [ctx.dot(psh.X, psh.XH), np.sqrt(ctx.dot(psh.Y, psh.Y)), ctx.diff_norm2(psh.X, psh.X0)] + ctx.kkt(psh.XH, psh.YH)
    assert tree == again
    assert np.allclose(tree, ered, rtol=1e-12, atol=1e-300)
    ctx.close()


@pytest.mark.parametrize("name", ["stocfor1", "rand300", "long_rows", "rand20k"])
@pytest.mark.parametrize("prec", [0, 1])
def test_fused_chunk_equals_unfused_ops(gpu_available, oracle, name, prec):
    """psh_chunk (2 fused kernels per iteration) against the op-by-op sequence, bit for bit,
    including the hat vectors written at the epoch start and on the last iteration."""
    from ps26119_amd import psh
    lp = model(name)
    K, tau, sigma, rho = 7, 0.3, 0.4, 1.0
    a = context(oracle, lp, prec)
    b = context(oracle, lp, prec)
    for ctx in (a, b):
        ctx.upload(psh.X, np.sin(1.0 + np.arange(lp.n)))
        ctx.upload(psh.Y, np.cos(2.0 + np.arange(lp.m)))
        ctx.copy(psh.X0, psh.X)
        ctx.copy(psh.Y0, psh.Y)
    for inner0 in (0, K):
        a.chunk(K, inner0, tau, sigma, rho)
        for s in range(K):
            w = (inner0 + s + 1.0) / (inner0 + s + 2.0)
            b.spmv_t(psh.Y, psh.ATY)
            b.r2h_primal(psh.X, psh.X0, psh.ATY, tau, w, rho, psh.XH, psh.XB)
            b.spmv(psh.XB, psh.AX)
            b.r2h_dual(psh.Y, psh.Y0, psh.AX, sigma, w, rho, psh.YH, psh.YB)
        for slot in (psh.X, psh.XH, psh.XB, psh.Y, psh.YH, psh.YB):
            assert np.array_equal(a.download(slot), b.download(slot)), (inner0, slot)
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["rand300", "long_rows", "rand20k"])
def test_phased_products_equal_row_tiles(gpu_available, oracle, name):
    """The phased half steps (persistent workgroups sweeping column slices in order, each row's
    running sum kept in CSR order) give the bits of the row-tiled kernels in every reduction mode;
    hip_sliced = 2 builds them at any size (the long-rows model has rows beyond the phased layout's
    limits and falls back to the tiles)."""
    from ps26119_amd import psh
    lp = model(name)
    K, tau, sigma, rho = 7, 0.3, 0.4, 1.0
    sc = psh.scaled_problem(oracle.scale(lp), lp)
    a = psh.Context(sc, sliced=2)
    b = psh.Context(sc, sliced=0)
    for mode in (psh.RED_TREE, psh.RED_CPU, psh.RED_DECISION):
        for ctx in (a, b):
            ctx.upload(psh.X, np.sin(1.0 + np.arange(lp.n)))
            ctx.upload(psh.Y, np.cos(2.0 + np.arange(lp.m)))
            ctx.copy(psh.X0, psh.X)
            ctx.copy(psh.Y0, psh.Y)
        for inner0 in (0, K):
            a.chunk(K, inner0, tau, sigma, rho, mode=mode)
            b.chunk(K, inner0, tau, sigma, rho, mode=mode)
        for slot in (psh.X, psh.XH, psh.XB, psh.Y, psh.YH, psh.YB):
            assert np.array_equal(a.download(slot), b.download(slot)), (mode, slot)
    a.close()
    b.close()


def _stats_bits(s):
    """Every field of a psh_stats, floats as their exact hex strings."""
    vals = [s.px, s.py, s.cross, s.r0_px, s.r0_py, s.r0_cross, s.have_r0, s.dx_norm, s.dy_norm, s.finite]
    vals += s.kkt.as_list()
    vals += [getattr(r, f) for r in (s.ray, s.drift) for f, _ in r._fields_]
    return [v.hex() if isinstance(v, float) else v for v in vals]


@pytest.mark.parametrize("name", MODELS)
def test_phased_check_equals_row_tiles(gpu_available, oracle, name):
    """The check pass over the phased layouts (the original values copied into the layouts' order;
    hip_sliced = 2 builds them at any size) gives the row tiles' bits: every scalar of the check
    in every reduction mode, with and without the ray and drift tests (the netlib models have
    unboxed columns, so the primal-ray products run too), the epoch residual, the stored Ã·Δx, and
    the KKT of named slots."""
    from ps26119_amd import psh
    lp = model(name)
    K, tau, sigma = 7, 0.3, 0.4
    sc = psh.scaled_problem(oracle.scale(lp), lp)
    a = psh.Context(sc, sliced=2)
    b = psh.Context(sc, sliced=0)
    for mode in (psh.RED_TREE, psh.RED_CPU, psh.RED_DECISION):
        for ctx in (a, b):
            ctx.upload(psh.X, np.sin(1.0 + np.arange(lp.n)))
            ctx.upload(psh.Y, np.cos(2.0 + np.arange(lp.m)))
            ctx.copy(psh.X0, psh.X)
            ctx.copy(psh.Y0, psh.Y)
            ctx.chunk(K, 0, tau, sigma, mode=mode)
        for flags in ((1, 1, 1), (0, 0, 1), (1, 0, 0)):
            sa, sb = a.check(1.3, 0.7, *flags, mode=mode), b.check(1.3, 0.7, *flags, mode=mode)
            assert _stats_bits(sa) == _stats_bits(sb), (mode, flags)
        assert np.array_equal(a.download(psh.AX), b.download(psh.AX)), mode
        assert a.kkt(psh.XH, psh.YH, mode) == b.kkt(psh.XH, psh.YH, mode), mode
    a.close()
    b.close()


def test_chunk_graphs_follow_the_reflection(gpu_available, oracle):
    """ρ is baked into a captured chunk's launches: chunks with ρ = 1, 0.5, 0.5, 1 on one context
    (graph replay) give the bits of the same chunks launched eagerly on another."""
    from ps26119_amd import psh
    lp = model("rand300")
    K, tau, sigma = 7, 0.3, 0.4
    a = context(oracle, lp, 0)
    b = context(oracle, lp, 0)
    b.graphs(0)
    for ctx in (a, b):
        ctx.upload(psh.X, np.sin(1.0 + np.arange(lp.n)))
        ctx.upload(psh.Y, np.cos(2.0 + np.arange(lp.m)))
        ctx.copy(psh.X0, psh.X)
        ctx.copy(psh.Y0, psh.Y)
    for k, rho in enumerate((1.0, 0.5, 0.5, 1.0)):
        for ctx in (a, b):
            ctx.chunk(K, K * k, tau, sigma, rho)
        for slot in (psh.X, psh.XB, psh.Y):
            assert np.array_equal(a.download(slot), b.download(slot)), (rho, slot)
    a.close()
    b.close()


_big = {}


def random_1m():
    from ps26119_amd import lpgen
    if "lp" not in _big:
        _big["lp"] = lpgen.random_lp(1_000_000, seed=1)
    return _big["lp"]


def test_full_size_chunk_is_the_oracle(gpu_available, oracle):
    """The 1,000,000-row random LP (6M nonzeros; phased products on both half steps): 64 fused
    iterations are the oracle's op sequence bit for bit at full size (x, x̂, y, ŷ, and the CPU-order
    reductions and KKT terms of that point) in CPU-order, tree and decision modes (the reductions
    only feed r⁰ inside a chunk)."""
    from ps26119_amd import psh
    lp = random_1m()
    sc = psh.scaled_problem(oracle.scale(lp), lp)
    exp, ered = oracle.backend_ops(lp, precision=0, iters=64, tau=0.3, sigma=0.4, rho=1.0)
    for mode in (psh.RED_CPU, psh.RED_TREE, psh.RED_DECISION):
        ctx = psh.Context(sc)
        ctx.upload(psh.X, np.sin(1.0 + np.arange(lp.n)))
        ctx.upload(psh.Y, np.cos(2.0 + np.arange(lp.m)))
        ctx.copy(psh.X0, psh.X)
        ctx.copy(psh.Y0, psh.Y)
        ctx.chunk(64, 0, 0.3, 0.4, 1.0, mode=mode)
        got = {"x": ctx.download(psh.X), "xh": ctx.download(psh.XH), "y": ctx.download(psh.Y),
               "yh": ctx.download(psh.YH)}
        for k in got:
            assert np.array_equal(got[k], exp[k]), (mode, k, float(np.max(np.abs(got[k] - exp[k]))))
        red = [ctx.dot(psh.X, psh.XH, psh.RED_CPU), np.sqrt(ctx.dot(psh.Y, psh.Y, psh.RED_CPU)),
               ctx.diff_norm2(psh.X, psh.X0, psh.RED_CPU)] + ctx.kkt(psh.XH, psh.YH, psh.RED_CPU)
        assert [float(v).hex() for v in red] == [float(v).hex() for v in ered]
        ctx.close()


def test_full_size_scaling_and_solve_prefix_are_the_oracle(gpu_available, oracle):
    """At 1,000,000 rows: psh_scale's factors are the oracle's bits, and a 128-iteration solve
    through the C API (presolve, device scaling, ‖Ã‖₂ with the reference's sequential Lanczos dots,
    two checks) ends on the oracle's x and y bits in CPU-order and default modes, so η, the scaled
    problem and the phased iterations are the reference's at full size."""
    import ps26119
    from ps26119_amd import psh
    lp = random_1m()
    sc = oracle.scale(lp)
    dev = psh.Context(psh.unscaled_problem(lp))
    rs, cs = dev.scale(12 if sc["geometric_mean_applied"] else 0, 10, 1)
    assert np.array_equal(rs, sc["row_scale"]) and np.array_equal(cs, sc["col_scale"])
    dev.close()
    e, v = oracle.solve(lp, presolve=True, iteration_limit=128)
    for prm in ({"hip_cpu_order": 1}, {}):
        r = ps26119.solve_model(lp, iteration_limit=128, params=prm)
        print(f"1M rows, {prm or 'default'}: {r.status_name} it={r.iterations} oracle it={e['iterations']}", flush=True)
        assert r.iterations == e["iterations"] == 128
        assert np.array_equal(r.x, v["x"]) and np.array_equal(r.y, v["y"]), prm


@pytest.mark.parametrize("name", ["stocfor1", "rand300", "rand20k"])
@pytest.mark.parametrize("mode", [0, 1])
def test_fused_check_equals_separate_reductions(gpu_available, oracle, name, mode):
    from ps26119_amd import psh
    lp = model(name)
    ctx = context(oracle, lp, 0)
    run_ops(ctx, lp, iters=5)
    st = ctx.check(omega=1.3, eta=0.9, want_ray=1, want_drift=1, want_pid=1, mode=mode)
    # residual pieces against the reference-interface ops
    ctx2 = context(oracle, lp, 0)
    run_ops(ctx2, lp, iters=5)
    ctx2.copy(psh.DX, psh.XB)
    ctx2.axpby(-1.0, psh.XH, 1.0, psh.DX)
    ctx2.copy(psh.DY, psh.YB)
    ctx2.axpby(-1.0, psh.YH, 1.0, psh.DY)
    ctx2.spmv(psh.DX, psh.AX)
    px, py, cr = ctx2.dot(psh.DX, psh.DX, mode), ctx2.dot(psh.DY, psh.DY, mode), ctx2.dot(psh.DY, psh.AX, mode)
    assert (st.px, st.py, st.cross) == (px, py, cr)
    assert st.dx_norm == ctx2.diff_norm2(psh.XH, psh.X0, mode)
    assert st.dy_norm == ctx2.diff_norm2(psh.YH, psh.Y0, mode)
    kk = ctx2.kkt(psh.XH, psh.YH, mode)
    assert [st.kkt.primal_residual, st.kkt.dual_residual, st.kkt.primal_obj, st.kkt.dual_obj] == kk[:4]
    assert st.finite == 1
    ctx.close()
    ctx2.close()


def test_kkt_cpu_order_matches_oracle_on_original_point(gpu_available, oracle):
    """KKT of the unscaled iterate, CPU order on the device vs the oracle's kkt_general."""
    from ps26119_amd import psh
    lp = model("rand20k")
    sc = oracle.scale(lp)
    ctx = context(oracle, lp, 0)
    run_ops(ctx, lp, iters=10)
    got = ctx.kkt(psh.XH, psh.YH, psh.RED_CPU)
    xs, ys = ctx.download(psh.XH), ctx.download(psh.YH)
    x = np.minimum(np.maximum(sc["col_scale"] * xs / sc["bound_scale"], lp.col_lower), lp.col_upper)
    y = sc["row_scale"] * ys / sc["obj_scale"]
    e = oracle.kkt(lp, x, y)
    exp = [e[k] for k in ("primal_residual", "dual_residual", "primal_obj", "dual_obj", "b_norm", "c_norm",
                          "primal_max_rel", "dual_max_rel")]
    assert [float(v).hex() for v in got] == [float(v).hex() for v in exp]
    ctx.close()


@pytest.mark.parametrize("name", MODELS + ["refinery"])
def test_device_scaling_is_bit_identical(gpu_available, oracle, name):
    """psh_scale (geometric mean when the entries span ≥ 10^4.5, Ruiz ×10, Pock–Chambolle) on the
    device gives the oracle's row and column factors bit for bit, and its scaled Ã, Ãᵀ give the
    same Ã·x and Ãᵀ·y (CSR-order sums) as a context loaded with the oracle's scaled values."""
    from ps26119_amd import lpgen, psh
    lp = lpgen.refinery_lp(T=6, seed=3) if name == "refinery" else model(name)
    sc = oracle.scale(lp)
    dev = psh.Context(psh.unscaled_problem(lp))
    rs, cs = dev.scale(12 if sc["geometric_mean_applied"] else 0, 10, 1)
    print(f"{name}: log10 range {sc['log10_range']:.2f}, geometric mean {sc['geometric_mean_applied']}", flush=True)
    assert np.array_equal(rs, sc["row_scale"]) and np.array_equal(cs, sc["col_scale"])
    ref = psh.Context(psh.scaled_problem(sc, lp))
    x, y = np.sin(1.0 + np.arange(lp.n)), np.cos(2.0 + np.arange(lp.m))
    for ctx in (dev, ref):
        ctx.upload(psh.X, x)
        ctx.upload(psh.Y, y)
        ctx.spmv(psh.X, psh.AX)
        ctx.spmv_t(psh.Y, psh.ATY)
    assert np.array_equal(dev.download(psh.AX), ref.download(psh.AX))
    assert np.array_equal(dev.download(psh.ATY), ref.download(psh.ATY))
    dev.close()
    ref.close()
