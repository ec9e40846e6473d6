"""Low-precision traversal at every instantiated row width.

The low-precision gather has one code path per padded width (BF16: one to
eight 64-byte k-blocks; FP8: coalesced 64-byte steps when the padded width is
a multiple of 64, operand-shaped loads otherwise), so each width is searched
here.

* BF16: values are integers in [0, 127], so every term of the norm-expansion
  distance (||q||^2 + ||x||^2 - 2 q.x, each at most 2 * 256 * 127^2) stays
  below 2^24 and BF16 must equal FP32 bit for bit.
* FP8: the returned (not re-ranked) distances must equal an independent model
  of the FP8 distance: c_j the mean of column j, rows coded as
  E4M3((x_j - c_j) / s_j) with s_j = max_i |x_ij - c_j| / 448, the query as
  E4M3((q_j - c_j) s_j / t) with t = max_j |(q_j - c_j) s_j| / 448 (nearest
  code, ties to the even code), distance ||q - c||^2 + ||x_hat||^2 - 2 t
  sum(codes products). Only the FP32 accumulation order differs, so the
  tolerance is relative 3e-5 of the terms' magnitude; a misplaced byte or a
  missing centre moves the distance by far more.
* Re-ranked FP8 results carry exact FP32 distances."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = [32, 64, 96, 160, 192, 200]


def half_scale(vmn, n, d, stream):
    v = vmn.datasets.integer_mixture(n, d, k=16, sigma=20.0, seed=77, stream=stream)
    return np.rint(v / 2).astype(np.float32)


def _e4m3_table():
    vals, codes = [], []
    for code in range(256):
        e, m = (code >> 3) & 0xF, code & 7
        if e == 15 and m == 7:
            continue  # NaN
        v = (m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7)
        vals.append(-v if code & 0x80 else v)
        codes.append(code)
    o = np.argsort(vals, kind="stable")
    return np.array(vals)[o], np.array(codes)[o]


def e4m3(x):
    """Nearest E4M3 value of each element of x (|x| <= 448), ties to the even code."""
    v, c = _e4m3_table()
    x = np.asarray(x, dtype=np.float64)
    hi = np.clip(np.searchsorted(v, x), 1, len(v) - 1)
    lo = hi - 1
    dlo, dhi = np.abs(x - v[lo]), np.abs(v[hi] - x)
    pick = np.where(dlo < dhi, lo, np.where(dhi < dlo, hi, np.where(c[lo] % 2 == 0, lo, hi)))
    return v[pick]


def fp8_model(base, q, ids):
    f32 = np.float32
    c = (base.astype(np.float64).sum(0) / len(base)).astype(f32)       # column means
    xb = (base - c).astype(f32)
    absmax = np.abs(xb).max(0).astype(f32)
    s = np.where(absmax > 0, absmax / f32(448.0), f32(1.0)).astype(f32)
    xc = e4m3(np.clip((xb / s).astype(f32), -448, 448))              # row codes
    xn = ((xc * s.astype(np.float64)) ** 2).sum(1)                      # ||x_hat||^2
    qq = (q - c).astype(f32)
    qs = (qq * s).astype(f32)
    mx = np.abs(qs).max(1)
    t = np.where(mx > 0, mx / f32(448.0), f32(1.0)).astype(f32)
    qc = e4m3(np.clip((qs / t[:, None]).astype(f32), -448, 448))       # query codes
    qn = (qq.astype(np.float64) ** 2).sum(1)
    dot = np.einsum("qd,qkd->qk", qc, xc[ids]) * t[:, None]
    mag = qn[:, None] + xn[ids] + 2 * np.abs(dot)
    return qn[:, None] + xn[ids] - 2 * dot, mag


@pytest.mark.parametrize("d", DIMS)
def test_lowp_search_at_width(vmn, gpu, d):
    base = half_scale(vmn, 3000, d, 0)
    q = half_scale(vmn, 64, d, 1)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2),
                     low_precision="both")
    fi, fd = idx.searchPoints(q, k=10, L=40)
    bi, bd = idx.searchPoints(q, k=10, L=40, prec="bf16")
    assert np.array_equal(fi, bi) and np.array_equal(fd, bd)
    ri, rd = idx.searchPoints(q, k=10, L=40, prec="bf16", rerank=20)
    assert np.array_equal(fi, ri) and np.array_equal(fd, rd)

    ei, ed = idx.searchPoints(q, k=10, L=40, prec="fp8")
    want, mag = fp8_model(base, q, ei)
    assert np.all(np.abs(ed - want) <= 3e-5 * mag + 1e-3), np.abs(ed - want).max()

    xi, xd = idx.searchPoints(q, k=10, L=40, prec="fp8", rerank=40)
    exact = ((base[xi].astype(np.float64) - q[:, None, :]) ** 2).sum(-1).astype(np.float32)
    assert np.array_equal(xd, exact)
    gt = np.argsort(((base[None, :, :].astype(np.float64) - q[:, None, :]) ** 2).sum(-1),
                    axis=1, kind="stable")[:, :10]
    assert vmn.recall_at_k(xi, gt) >= vmn.recall_at_k(fi, gt) - 0.05


@pytest.mark.parametrize("d", DIMS)
@pytest.mark.parametrize("mode", ["fresh", "parth"])
def test_exact_bf16_build_at_width(vmn, orc, gpu, d, mode):
    """On rows exact in BF16 the build's searches and prunes (the 16-row MFMA
    Gram) read BF16 rows; the graph must be the oracle's FP32 graph at every
    padded width (one to eight 64-byte k-blocks per row). R = 12 stops most
    prunes at R, with candidates left over."""
    base = half_scale(vmn, 1500, d, 0)
    if mode == "fresh":
        idx = vmn.Vamana(base, mode="fresh", R=12, L_build=40, alphas=(1.0, 1.2))
        oadj, med = orc.fresh_build(base, R=12, L=40, alphas=(1.0, 1.2))
    else:
        idx = vmn.Vamana(base, mode="parth", R=12, L_build=40, alphas=(1.2, 1.2))
        oadj, med = orc.parth_build(base, R=12, L=40, alphas=(1.2, 1.2))
    assert idx.build_precision == "bf16"
    assert idx.medoid == med
    assert np.array_equal(idx.adjacency, oadj)
    idx.close()


@pytest.mark.parametrize("d", [32, 96, 200])
def test_exact_bf16_consolidation_at_width(vmn, orc, gpu, d):
    """Delete consolidation on exact BF16 rows (occlusion checks as 16-row
    MFMA Gram masks, then the MFMA prune from the kept members on) equals the
    oracle's FP32 consolidation at other padded widths than 128."""
    base = half_scale(vmn, 1500, d, 0)
    idx = vmn.Vamana(base, mode="fresh", R=16, L_build=40, alphas=(1.0, 1.2))
    adj, med = orc.fresh_build(base, R=16, L=40, alphas=(1.0, 1.2))
    assert idx.build_precision == "bf16" and np.array_equal(idx.adjacency, adj)
    kill = np.random.default_rng(d).choice(1500, 300, replace=False).astype(np.int32)
    dele = np.zeros(1500, np.uint8)
    dele[kill] = 1
    idx.deleteIds(kill)
    idx.consolidate()
    orc.consolidate(base, adj, dele, alpha=1.2)
    assert np.array_equal(idx.adjacency, adj)
    idx.close()
