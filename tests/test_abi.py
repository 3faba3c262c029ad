"""CPU-side checks of the C ABI: the library loads, exports every symbol the
public headers declare, the ctypes struct layout equals the C layout, and
argument validation returns the reference's status codes without touching
a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import block_ipt_amd
from block_ipt_amd import abi, api


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                             text, flags=re.M):
            if not m.group(1).startswith(("ipt", "acx")):
                continue
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    l = abi.lib()
    names = _declared_functions()
    assert "ipt_cuda_status_string" in names and "ipt_cuda_sparse_csc_device_tol" in names
    missing = [n for n in sorted(names) if not hasattr(l, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared surface
    assert set(abi.SIGNATURES) == names


def test_result_struct_layout_matches_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "ipt_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(IPTCudaResult),
         offsetof(IPTCudaResult, matvecs), offsetof(IPTCudaResult, davidson_history),
         offsetof(IPTCudaResult, best_so_far_source), offsetof(IPTCudaResult, cluster_soft_locked_count),
         sizeof(IPTDavidsonHistoryEntry), sizeof(IPTDavidsonBlockHistoryEntry));
  return 0; }
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    R = abi.IPTCudaResult
    want = [ctypes.sizeof(R), R.matvecs.offset, R.davidson_history.offset, R.best_so_far_source.offset,
            R.cluster_soft_locked_count.offset, ctypes.sizeof(abi.IPTDavidsonHistoryEntry),
            ctypes.sizeof(abi.IPTDavidsonBlockHistoryEntry)]
    assert got == want


def test_status_strings():
    l = abi.lib()
    assert l.ipt_cuda_status_string(0) == b"success"
    assert l.ipt_cuda_status_string(1) == b"invalid argument"
    assert l.ipt_cuda_status_string(2) == b"allocation failed"
    assert l.ipt_cuda_status_string(99) == b"unknown error"


def test_free_result_resets_fields():
    r = abi.IPTCudaResult()
    r.n = 5
    abi.lib().ipt_cuda_free_result(ctypes.byref(r))
    assert r.n == 0 and not r.vectors and np.isnan(r.fixed_point_residual)
    assert r.davidson_target_index == -1 and r.pair28_lock_state == b"unknown"


@pytest.mark.parametrize("n,k,tol,maxiter", [(0, 1, 1e-12, 10), (4, 0, 1e-12, 10), (4, 5, 1e-12, 10),
                                             (4, 2, 0.0, 10), (4, 2, 1e-12, 0), (4, 2, -1.0, 10)])
def test_invalid_arguments_rejected_before_gpu(n, k, tol, maxiter):
    H = np.eye(max(n, 1))
    with pytest.raises(abi.IptError) as e:
        api.ipt_cuda_tol(H[:max(n, 1), :max(n, 1)] if n else H, k, tol, maxiter) if n else \
            _raw_tol(n, k, tol, maxiter)
    assert e.value.status == 1


def _raw_tol(n, k, tol, maxiter):
    r = abi.IPTCudaResult()
    H = np.eye(2)
    st = abi.lib().ipt_cuda_tol(H.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, k, tol, maxiter,
                                ctypes.byref(r))
    raise abi.IptError(st, "ipt_cuda_tol")


def test_null_pointers_rejected():
    r = abi.IPTCudaResult()
    assert abi.lib().ipt_cuda(None, 4, 2, 5, ctypes.byref(r)) == 1
    assert abi.lib().ipt_cuda_sparse_csc(None, None, None, 4, 2, 0, 5, ctypes.byref(r)) == 1
    dv, dl = ctypes.c_void_p(), ctypes.c_void_p()
    assert abi.lib().ipt_cuda_device(None, 4, 2, 5, None, ctypes.byref(dv), ctypes.byref(dl)) == 1
    assert dv.value is None and dl.value is None


def test_mixed_result_struct_layout_matches_c(tmp_path):
    src = tmp_path / "probe_mixed.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "ipt_hip_mixed.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n",
         sizeof(IPTMixedCudaDeviceResult), offsetof(IPTMixedCudaDeviceResult, fp16_time_sec),
         offsetof(IPTMixedCudaDeviceResult, d_vectors), sizeof(IPTMixedFP32CudaDeviceResult),
         sizeof(IPTMixedTF32CudaDeviceResult), sizeof(IPTRayleighMixedCudaDeviceResult),
         offsetof(IPTRayleighMixedCudaDeviceResult, fp64_fixed_point_residual),
         offsetof(IPTRayleighMixedCudaDeviceResult, d_values));
  return 0; }
""")
    exe = tmp_path / "probe_mixed"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    M, R = abi.IPTMixedCudaDeviceResult, abi.IPTRayleighMixedCudaDeviceResult
    want = [ctypes.sizeof(M), M.fp16_time_sec.offset, M.d_vectors.offset,
            ctypes.sizeof(abi.IPTMixedFP32CudaDeviceResult), ctypes.sizeof(abi.IPTMixedTF32CudaDeviceResult),
            ctypes.sizeof(R), R.fp64_fixed_point_residual.offset, R.d_values.offset]
    assert got == want


def test_mixed_free_resets_fields():
    r = abi.IPTMixedTF32CudaDeviceResult()
    r.n = 7
    r.tf32_iterations = 3
    abi.lib().ipt_mixed_precision_tf32_cuda_free_device_result(ctypes.byref(r))
    assert r.n == 0 and r.tf32_iterations == 0 and not r.d_vectors
    assert np.isnan(r.tf32_time_sec) and np.isnan(r.fp64_fixed_point_residual)


def test_mixed_invalid_arguments_rejected_before_gpu():
    r = abi.IPTMixedCudaDeviceResult()
    st = abi.lib().ipt_mixed_precision_cuda_device_tol(None, 4, 2, 1e-3, 10, 1e-12, 10, None, ctypes.byref(r))
    assert st == 1 and not r.d_vectors
    st = abi.lib().ipt_mixed_precision_cuda_device_tol(ctypes.c_void_p(8), 4, 2, 0.0, 10, 1e-12, 10, None,
                                                       ctypes.byref(r))
    assert st == 1


def test_acx_result_and_arguments_before_gpu(tmp_path):
    """ACXCudaResult layout equals the C layout; reset gives NaN residuals;
    invalid arguments return status 1 before any device call
    (IPT/src/ACX/acx_cuda.cu:420-432)."""
    src = tmp_path / "probe.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "ipt_hip_acx.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(ACXCudaResult), offsetof(ACXCudaResult, time_sec),
         offsetof(ACXCudaResult, fixed_point_residual), offsetof(ACXCudaResult, d_values));
  return 0; }
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    R = abi.ACXCudaResult
    assert got == [ctypes.sizeof(R), R.time_sec.offset, R.fixed_point_residual.offset, R.d_values.offset]
    l = abi.lib()
    assert l.acx_cuda_status_string(5) == b"zero diagonal denominator"
    assert l.acx_cuda_status_string(0) == b"success"
    r = R()
    r.iterations = 3
    l.acx_cuda_reset_result(ctypes.byref(r))
    assert r.iterations == 0 and np.isnan(r.max_residual) and np.isnan(r.fixed_point_residual)
    p = ctypes.c_void_p(8)
    for args in [(None, 4, 2, 1e-12, 10), (p, 0, 1, 1e-12, 10), (p, 4, 0, 1e-12, 10), (p, 4, 5, 1e-12, 10),
                 (p, 4, 2, 0.0, 10), (p, 4, 2, 1e-12, 0)]:
        r.iterations = 7
        assert l.acx_cuda_device_tol(*args, None, ctypes.byref(r)) == 1
        assert r.iterations == 0 and not r.d_vectors
    assert l.acx_cuda_device_tol(p, 4, 2, 1e-12, 10, None, None) == 1


def _lines_brute(w, ld, elem, rows=4096):
    """128-byte lines touched by rows 0..rows-1 of a row-major array of w
    elements per row at stride ld, counted row by row."""
    total = 0
    for r in range(rows):
        first = (r * ld * elem) // 128
        last = (r * ld * elem + w * elem - 1) // 128
        total += last - first + 1
    return total / rows


def test_csr_row_stride_layout():
    """The row stride the block path gives X before a CSR product: never
    more lines per row than the unpadded stride (brute-force count over
    4096 rows), aligned widths kept, within the block path's allocation
    bound (w rounded up to a line), and the measured cases (8-byte rows:
    12 -> 16, 27 -> 32; 20 and 40 already touch the fewest lines)."""
    f = abi.lib().ipt_hip_csr_row_stride
    assert f(0, 8) == -1 and f(5, 2) == -1
    for elem in (8, 4):
        per = 128 // elem
        for w in range(1, 131):
            ld = f(w, elem)
            assert w <= ld <= -(-w // per) * per
            assert ld in (w, 1 << (w - 1).bit_length(), -(-w // per) * per)
            assert _lines_brute(w, ld, elem) <= _lines_brute(w, w, elem) + 1e-12
            if (w * elem) % 128 == 0:
                assert ld == w
            # no candidate touches fewer lines
            for c in {w, 1 << (w - 1).bit_length(), -(-w // per) * per}:
                assert _lines_brute(w, ld, elem) <= _lines_brute(w, c, elem) + 1e-12
    assert [f(w, 8) for w in (12, 27, 20, 40)] == [16, 32, 20, 40]
    assert f(27, 4) == 32 and f(1, 8) == 1
