"""fixResult + printResult's row selection on the device
(clast_result_finalize) against a row-by-row walk of printResult
(CHostResultHolder.cu:131-243): the last printed row's ten fields, a count
per query label reset when the label changes from the last printed one,
the walk starting from the empty label."""
from __future__ import annotations

import numpy as np
import pytest

from clast_amd import native

pytestmark = pytest.mark.gpu


def walk(rows, qlabel, tlabel, tstart, num_out):
    """printResult, one row at a time, over rows already in fixResult order."""
    pre = None
    pre_label = 0  # the empty label
    count = 0
    out = []
    for i in range(len(rows["qID"])):
        lab = qlabel[i]
        if lab != pre_label:
            count = 0
        if not (num_out == -1 or lab != pre_label or count < num_out):
            out.append(False)
            continue
        f = (lab, rows["qIdx"][i], rows["qID"][i] & 1, tlabel[i], rows["tIdx"][i] + tstart[i], rows["tHL"][i],
             rows["qHL"][i], rows["match"][i], rows["score"][i], rows["evalue"][i])
        if pre is not None and f == pre:
            out.append(False)
            continue
        out.append(True)
        count += 1
        pre, pre_label = f, lab
    return np.array(out, dtype=bool)


@pytest.mark.parametrize("num_out", [-1, 0, 1, 2, 5])
def test_finalize_matches_the_walk(num_out):
    native.init(0)
    rng = np.random.default_rng(50 + num_out)
    n = 4000
    # few distinct values so that equal rows and equal labels are common;
    # query label ids repeat across reads and include the empty label (0)
    rows = {"tID": rng.integers(0, 5, n), "tIdx": rng.integers(0, 3, n), "qID": rng.integers(0, 60, n),
            "qIdx": rng.integers(0, 2, n), "tHL": np.full(n, 150), "qHL": np.full(n, 150),
            "match": rng.integers(148, 150, n), "score": rng.integers(140, 143, n)}
    rows = {k: v.astype(np.int32) for k, v in rows.items()}
    rows["evalue"] = 0.71 * 1e6 * 150 * np.exp(-1.37 * rows["score"].astype(np.float64))
    read_label = rng.integers(0, 12, 30)  # 30 reads, 12 label ids (0 = empty)
    qlabel = read_label[rows["qID"] // 2].astype(np.int32)
    tlabel = (rows["tID"] % 3).astype(np.int32)  # targets 0/3 and 1/4 share labels
    tstart = np.where(rows["tID"] == 3, -1, 0).astype(np.int32)  # tID 3 at tIdx + (-1) ~ tID 0 at tIdx - 1
    order, printed, npr = native.result_finalize(rows, qlabel, tlabel, tstart, num_out)
    # fixResult: stable (qID / 2 ascending, score descending)
    exp_order = np.lexsort((-rows["score"].astype(np.int64), rows["qID"] // 2))
    np.testing.assert_array_equal(order, exp_order)
    srt = {k: v[exp_order] for k, v in rows.items()}
    exp = walk(srt, qlabel[exp_order].tolist(), tlabel[exp_order].tolist(), tstart[exp_order].tolist(), num_out)
    np.testing.assert_array_equal(printed.astype(bool), exp)
    assert npr == int(exp.sum()) and 0 < npr < n
