"""airan_ldpc_decode_EnqueueNoAlloc (the decoder with its message state in
the library's fixed device array, for callers that cannot allocate) against
airan_ldpc_decode_Enqueue with a caller workspace: identical bits, clamped
posteriors and iteration counts on two calls in a row, the second with fewer
code blocks (the array keeps the first call's state, which the decoder must
not read); requests beyond the array are rejected before anything runs."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import ptr
from puschrx import abi, ops

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

Z = 384


def _llrs(oracle, ncb, snr_db, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncb):
        info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
        cw = np.zeros(68 * Z, np.uint8)
        oracle.or_ldpc_encode(Z, ptr(info), ptr(cw))
        s2 = 10 ** (-snr_db / 10)
        y = (1.0 - 2.0 * cw) + rng.normal(0, np.sqrt(s2), cw.size)
        llr = np.clip(np.rint(2 * y / s2 * 256), -5120, 5120).astype(np.int16)
        llr[: 2 * Z] = 0
        out.append(llr)
    return np.concatenate(out)


def _args(lam, bits, lam_out, iters, ncb, rows, mode, iterations, early):
    a = abi.LdpcDecodeOpArgsV1()
    a.abi_version, a.struct_size = abi.ABI_VERSION, C.sizeof(a)
    a.lam_in, a.bits = lam.data_ptr(), bits.data_ptr()
    a.lam_out, a.iterations_out = lam_out.data_ptr(), iters.data_ptr()
    a.num_cb, a.z, a.k, a.max_iterations = ncb, Z, 22 * Z, iterations
    a.active_rows, a.mode, a.early_stop = rows, mode, 1 if early else 0
    a.stream = torch.cuda.current_stream().cuda_stream
    return a


def _decode(lib, lam, ncb, rows, mode, iterations, early, no_alloc):
    bits = torch.full((ncb * 22 * Z,), 7, dtype=torch.int8, device="cuda")
    lam_out = torch.zeros(ncb * 68 * Z, dtype=torch.int16, device="cuda")
    iters = torch.zeros(ncb, dtype=torch.uint8, device="cuda")
    a = _args(lam, bits, lam_out, iters, ncb, rows, mode, iterations, early)
    if no_alloc:
        st = lib.airan_ldpc_decode_EnqueueNoAlloc(C.byref(a))
    else:
        nws = lib.airan_ldpc_decode_WorkspaceBytes(ncb, Z, rows)
        ws = torch.empty(max(nws, 8), dtype=torch.uint8, device="cuda")
        st = lib.airan_ldpc_decode_Enqueue(C.byref(a), C.c_void_p(ws.data_ptr()), ws.numel())
    assert st == abi.OK
    torch.cuda.synchronize()
    return bits.cpu().numpy(), lam_out.cpu().numpy(), iters.cpu().numpy()


@pytest.mark.parametrize("mode,rows,iterations,early", [(abi.LDPC_MODE_REF, 46, 3, False),
                                                        (abi.LDPC_MODE_FP16, 46, 8, True),
                                                        (abi.LDPC_MODE_FP16, 6, 8, True)],
                         ids=["ref46", "fp16_46", "fp16_6"])
def test_no_alloc_decode_matches_workspace_decode(oracle, gpu_stream, mode, rows, iterations, early):
    lib = abi.load_library()
    assert lib.airan_InitDevice() == abi.OK
    for call, (ncb, snr, seed) in enumerate(((40, 2.0, 1), (7, 1.5, 2))):
        print(f"no-alloc decode: mode {mode} rows {rows} call {call} ({ncb} code blocks)", flush=True)
        lam = torch.from_numpy(_llrs(oracle, ncb, snr, seed)).cuda()
        want = _decode(lib, lam, ncb, rows, mode, iterations, early, no_alloc=False)
        got = _decode(lib, lam, ncb, rows, mode, iterations, early, no_alloc=True)
        for w, g in zip(want, got):
            assert np.array_equal(w, g)
        assert set(np.unique(got[0])) <= {0, 1}


def test_no_alloc_rejects_beyond_the_array(gpu_stream):
    lib = abi.load_library()
    cap = lib.airan_ldpc_decode_NoAllocMaxWorkspaceBytes()
    assert cap == 32 << 20
    per_cb = lib.airan_ldpc_decode_WorkspaceBytes(1, Z, 46)
    ncb = cap // per_cb + 1
    assert lib.airan_ldpc_decode_WorkspaceBytes(ncb, Z, 46) > cap
    assert lib.airan_ldpc_decode_WorkspaceBytes(143, Z, 46) <= cap  # the reference's 23-slot TB fits
    lam = torch.zeros(ncb * 68 * Z, dtype=torch.int16, device="cuda")
    bits = torch.full((ncb * 22 * Z,), 7, dtype=torch.int8, device="cuda")
    lam_out = torch.zeros_like(lam)
    iters = torch.zeros(ncb, dtype=torch.uint8, device="cuda")
    a = _args(lam, bits, lam_out, iters, ncb, 46, abi.LDPC_MODE_REF, 3, False)
    assert lib.airan_ldpc_decode_EnqueueNoAlloc(C.byref(a)) == abi.RESOURCE_TOO_SMALL
    torch.cuda.synchronize()
    assert int(bits[0].item()) == 7  # nothing ran
