"""CPU checks of the HIP library's host side (no kernel launches)."""
import ctypes
import os
import re

import numpy as np

import edlut_amd as E
from edlut_amd import _capi
from tests.helpers import np_out_csr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "edlut_amd.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s+\**\s*(edlut_[a-z_0-9]+)\s*\(", header, re.M))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} is declared but not exported"
    assert declared == set(_capi.EXPORTS)


def test_csr_builder_matches_reference_sort_order():
    """edlut_build_out_csr == an independent restatement of qsort_inters."""
    rng = np.random.default_rng(0)
    n_src, n_tgt, D, e = 300, 200, 7, 20000
    src = rng.integers(0, n_src, e)
    tgt = rng.integers(0, n_tgt, e)
    # no exact duplicates: their relative order is unspecified in the reference
    key = (src * n_tgt + tgt) * D
    _, keep = np.unique(key, return_index=True)
    src, tgt = src[keep], tgt[keep]
    d = rng.integers(1, D + 1, len(src))
    typ = rng.integers(0, 3, len(src)).astype(np.uint8)
    w = rng.random(len(src)).astype(np.float32)
    a = E.build_out_csr(n_src, n_tgt, D, src, tgt, d, typ, w)
    b = np_out_csr(n_src, n_tgt, D, src, tgt, d, typ, w)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_csr_builder_rejects_bad_edges():
    import pytest

    with pytest.raises(E.EDLUTException):
        E.build_out_csr(10, 10, 3, [0], [0], [4], [0], [1.0])  # delay > max_delay
    with pytest.raises(E.EDLUTException):
        E.build_out_csr(10, 10, 3, [0], [10], [1], [0], [1.0])  # target out of range
    with pytest.raises(E.EDLUTException):
        E.build_out_csr(10, 10, 3, [0], [1], [1], [3], [1.0])  # type 3 takes currents


def test_drive_hash_host_copy_matches_oracle(oracle):
    lib = _capi.lib()
    for seed, step, i in [(0, 1, 0), (123456789, 77, 999999), (2**63, 10**6, 2**30 - 1)]:
        assert lib.edlut_drive_hash(seed, step, i) == oracle.drive_hash(seed, step, i)


def test_binomial_thresholds_distribution():
    """k = #{j: u < thr[j]} for uniform u reproduces Binomial(100, 5e-4)."""
    from math import comb

    thr = E.binomial_thresholds(100, 5e-4).astype(np.float64) / 2.0**32
    p = 5e-4
    pmf = [comb(100, k) * p**k * (1 - p) ** (100 - k) for k in range(8)]
    tails = [sum(pmf[j:]) for j in range(1, len(thr) + 1)]
    assert np.allclose(thr, tails, rtol=0, atol=2**-32)
    assert np.all(np.diff(thr) <= 0)


def test_build_flags_keep_packed_fp32_out(tmp_path):
    """The library's hipcc flags drop packed FP32 instructions from device
    code: a kernel whose float pairs hipcc -O3 would otherwise fuse into
    v_pk_fma_f32 compiles to none under them (the warning hipcc prints about
    the feature comes from its host pass)."""
    import shutil
    import subprocess

    from edlut_amd import build
    if not os.path.exists(build.HIPCC) and not shutil.which("hipcc"):
        import pytest
        pytest.skip("hipcc not available")
    src = tmp_path / "pk.hip"
    src.write_text("#include <hip/hip_runtime.h>\n"
                   "__global__ void k(float2 *a, const float2 *b, int n) {\n"
                   "  int i = blockIdx.x * blockDim.x + threadIdx.x;\n"
                   "  if (i < n) { float2 x = a[i], y = b[i]; x.x = x.x * y.x + y.y; x.y = x.y * y.y + y.x; a[i] = x; }\n"
                   "}\n")

    def device_asm(flags):
        out = tmp_path / "pk.s"
        r = subprocess.run([build.HIPCC, "--offload-arch=gfx950", "-O3", "--cuda-device-only", "-S", "-o", str(out),
                            str(src), *flags], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        return out.read_text()

    lib_flags = [f for f in build.FLAGS if f not in ("-shared", "-fPIC") and not f.startswith("--offload-arch")]
    lib_flags = [f for f in lib_flags if f != "-I" and not f.endswith("include")]
    assert "v_pk_fma_f32" in device_asm(["-ffp-contract=fast"])       # the pattern does fuse without the flag
    assert not re.search(r"v_pk_(fma|add|mul)_f32", device_asm(lib_flags + ["-ffp-contract=fast"]))
