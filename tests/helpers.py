"""Shared helpers of the test suite."""
from __future__ import annotations

import numpy as np

ADEX_DEFAULTS = np.array([1.0, 9.0, 2.0, -50.0, 50.0, 0.0, -80.0, -80.0, -65.0, 10.0, 110.0, 5.0, 10.0, 20.0],
                         np.float32)
LIF_DEFAULTS = np.array([0.0, -80.0, -65.0, -50.0, 2.0, 5.0, 10.0, 1.0, 0.2, 20.0], np.float32)
HH_DEFAULTS = np.array([0.0, -80.0, -65.0, 10.0, 120.0, -52.0, 5.0, 10.0, 20.0, 20000.0, 6000.0, 50.0, -90.0],
                       np.float32)


GRANULE_DEFAULTS = np.zeros(0, np.float32)
PURKINJE_DEFAULTS = np.array([0.0, -80.0, -35.0, -70.0, 1.0, 2.0, 20.0, 1.35], np.float32)
IZH_DEFAULTS = np.array([0.1, 0.23, -65.0, 0.2, 0.0, -80.0, 100.0, 5.0, 10.0, 20.0], np.float32)
LIF_IS_DEFAULTS = np.array([0.0, -80.0e-3, -65.0e-3, -50.0e-3, 2.0e-12, 5.0e-3, 10.0e-3, 1.0e-3, 0.2e-9, 20.0e-3],
                           np.float32)


def defaults(model: str) -> np.ndarray:
    return {"AdExTimeDrivenModel_GPU": ADEX_DEFAULTS, "LIFTimeDrivenModel_GPU": LIF_DEFAULTS,
            "HHTimeDrivenModel_GPU": HH_DEFAULTS, "EgidioGranuleCell_TimeDriven_GPU": GRANULE_DEFAULTS,
            "TimeDrivenPurkinjeCell_GPU": PURKINJE_DEFAULTS, "IzhikevichTimeDrivenModel_GPU": IZH_DEFAULTS,
            "LIFTimeDrivenModel_IS_GPU": LIF_IS_DEFAULTS}[model]


def np_out_csr(n_src, n_tgt, max_delay, src, tgt, delay, typ, w):
    """Independent numpy restatement of qsort_inters (Kernel/src/spike/Network.cpp:45-68)
    with one queue: order by source, then delay, then target (stable)."""
    src = np.asarray(src, np.int64)
    tgt = np.asarray(tgt, np.int64)
    delay = np.asarray(delay, np.int64)
    order = np.lexsort((tgt, delay, src))
    row = src[order] * max_delay + (delay[order] - 1)
    row_ptr = np.zeros(n_src * max_delay + 1, np.int64)
    np.add.at(row_ptr, row + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    tt = (tgt[order].astype(np.uint32) | (np.asarray(typ, np.uint32)[order] << np.uint32(30))).astype(np.uint32)
    return row_ptr, tt, np.asarray(w, np.float32)[order]


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)
