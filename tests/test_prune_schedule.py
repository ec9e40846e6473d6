"""The 16-row schedule of the MFMA prune (robust_prune_mfma in
cs5013-project_amd/csrc/vmn_prune.hpp) against the sequential robust prune
(parth/src/build_graph.cu:241-258 strict, vamana/include/kernels.cuh:329-353
with <=), on small integer lists full of equal distances. Both take the same
exact distances; what is checked is the order of decisions: speculative p*
rows taken in list order, skipped when an earlier p* dropped them, masks
applied only to later candidates, the scan resumed after the 16th row, and the
stop at R. The planted mistakes below (a dropped row still emitted or still
applying its mask, the scan resumed at the wrong place) each change some
output here. The GPU kernel itself is checked against the oracle's built
graphs in the GPU tests."""
import numpy as np
import pytest

ROWS = 16


def sequential(D, cd, alpha, R, strict):
    cnt, kept, out, frm = len(cd), [True] * len(cd), [], 0
    a = np.float32(alpha)
    while len(out) < R:
        p = next((j for j in range(frm, cnt) if kept[j]), -1)
        if p < 0:
            break
        out.append(p)
        if len(out) >= R:
            break
        for j in range(p + 1, cnt):
            if kept[j]:
                lhs = a * D[p, j]
                if (lhs < cd[j]) if strict else (lhs <= cd[j]):
                    kept[j] = False
        frm = p + 1
    return out


def sixteen_rows(D, cd, alpha, R, strict):
    """The kernel's schedule: masks for 16 rows from one pass, then applied."""
    cnt, kept, out, frm = len(cd), [True] * len(cd), [], 0
    a = np.float32(alpha)
    while len(out) < R:
        live = [j for j in range(frm, cnt) if kept[j]]
        if not live:
            break
        p0, act = live[0], live[1:]
        ns = min(len(live), ROWS)
        rows = [p0] + act[: ROWS - 1]
        pm = [0] * len(act)
        for k, c in enumerate(act):
            for j in range(ns):
                lhs = a * D[rows[j], c]
                pr = (lhs < cd[c]) if strict else (lhs <= cd[c])
                if pr and k >= j:
                    pm[k] |= 1 << j
        keep = [True] * len(act)
        done = False
        for j in range(ns):
            if j > 0 and not keep[j - 1]:
                continue
            out.append(rows[j])
            if len(out) >= R:
                done = True
                break
            for k in range(len(act)):
                if (pm[k] >> j) & 1:
                    keep[k] = False
        if done:
            break
        for k, c in enumerate(act):
            if not keep[k]:
                kept[c] = False
        frm = (p0 if ns == 1 else act[ns - 2]) + 1
    return out


@pytest.mark.parametrize("seed", range(6))
def test_sixteen_row_schedule_equals_sequential(seed):
    rng = np.random.default_rng(seed)
    for _ in range(25):
        n = int(rng.integers(1, 160))
        d = int(rng.integers(1, 5))
        X = rng.integers(0, 4, size=(n + 1, d)).astype(np.float32)
        C, q = X[:-1], X[-1]
        cd = ((C - q) ** 2).sum(1).astype(np.float32)
        order = np.argsort(cd, kind="stable")
        C, cd = C[order], cd[order]
        D = ((C[:, None, :] - C[None, :, :]) ** 2).sum(2).astype(np.float32)
        alpha = float(rng.choice([1.0, 1.2, 1.5]))
        R = int(rng.integers(1, 70))
        strict = bool(rng.integers(0, 2))
        assert sixteen_rows(D, cd, alpha, R, strict) == sequential(D, cd, alpha, R, strict)


@pytest.mark.parametrize("mistake", ["emit_dropped", "resume_after_p0", "mask_before_emit"])
def test_schedule_mistakes_are_caught(mistake):
    """Each planted schedule mistake changes the output on some of these lists:
    emitting a row an earlier p* dropped, resuming the scan just after p0
    (rows already output come back), or applying a row's mask before checking
    whether it is still kept."""
    rng = np.random.default_rng(11)

    def planted(D, cd, alpha, R, strict):
        cnt, kept, out, frm = len(cd), [True] * len(cd), [], 0
        a = np.float32(alpha)
        while len(out) < R:
            live = [j for j in range(frm, cnt) if kept[j]]
            if not live:
                break
            p0, act = live[0], live[1:]
            ns = min(len(live), ROWS)
            rows = [p0] + act[: ROWS - 1]
            pm = [0] * len(act)
            for k, c in enumerate(act):
                for j in range(ns):
                    lhs = a * D[rows[j], c]
                    if ((lhs < cd[c]) if strict else (lhs <= cd[c])) and k >= j:
                        pm[k] |= 1 << j
            keep = [True] * len(act)
            done = False
            for j in range(ns):
                if mistake == "mask_before_emit":
                    for k in range(len(act)):
                        if (pm[k] >> j) & 1:
                            keep[k] = False
                if j > 0 and not keep[j - 1] and mistake != "emit_dropped":
                    continue
                out.append(rows[j])
                if len(out) >= R:
                    done = True
                    break
                if mistake != "mask_before_emit":
                    for k in range(len(act)):
                        if (pm[k] >> j) & 1:
                            keep[k] = False
            if done:
                break
            for k, c in enumerate(act):
                if not keep[k]:
                    kept[c] = False
            frm = p0 + 1 if mistake == "resume_after_p0" else (p0 if ns == 1 else act[ns - 2]) + 1
        return out

    differs = 0
    for _ in range(40):
        n = int(rng.integers(20, 160))
        X = rng.integers(0, 4, size=(n + 1, 3)).astype(np.float32)
        C, q = X[:-1], X[-1]
        cd = ((C - q) ** 2).sum(1).astype(np.float32)
        order = np.argsort(cd, kind="stable")
        C, cd = C[order], cd[order]
        D = ((C[:, None, :] - C[None, :, :]) ** 2).sum(2).astype(np.float32)
        differs += planted(D, cd, 1.2, 64, False) != sequential(D, cd, 1.2, 64, False)
    assert differs > 0
