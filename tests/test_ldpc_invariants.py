"""Invariants of the FP16-message layered min-sum decoder that the GPU decoder
relies on, checked on the oracle decoder (whose fp16 trajectory the GPU
decoder reproduces bit for bit, tests/test_gpu_stages.py):

* no posterior is ever -0: the inputs are +0 or non-zero and round-to-nearest
  sums and differences give -0 only from -0 operands, so a hard decision
  (negative, non-zero, not NaN) is the posterior's sign bit and a syndrome
  row's parity is the XOR of its posteriors' sign bits;
* the posteriors stay finite and far below the fp16 range for every int16
  input (every row holds a parity column of degree one or two, whose
  extrinsic caps the row's other messages at 0.75 of it), so no extrinsic is
  ever infinite or NaN.

The inputs are the extremes: full-range random int16 LLRs, clean codewords at
the largest int16 magnitudes (the fastest growth), and +-5120 noise; up to the
decoder's iteration limit (100) without early stop.
"""
import numpy as np
import pytest

from oracle_lib import ptr

Z, NCB = 384, 2
FP16 = 1


def _inputs(oracle, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "full_int16":
        lam = rng.integers(-32768, 32768, NCB * 68 * Z).astype(np.int16)
    elif kind == "noise_5120":
        lam = rng.choice(np.array([-5120, 5120], np.int16), NCB * 68 * Z)
    else:
        blocks = []
        for _ in range(NCB):
            info = rng.integers(0, 2, 22 * Z, dtype=np.uint8)
            cw = np.zeros(68 * Z, np.uint8)
            oracle.or_ldpc_encode(Z, ptr(info), ptr(cw))
            blocks.append(np.where(cw == 0, 32767, -32768).astype(np.int16))
        lam = np.concatenate(blocks)
    lam = lam.reshape(NCB, 68 * Z)
    lam[:, : 2 * Z] = 0
    return lam.reshape(-1).copy()


@pytest.mark.parametrize("kind", ["full_int16", "noise_5120", "clean_max"])
@pytest.mark.parametrize("rows,iters", [(4, 100), (6, 100), (8, 100), (46, 60)])
def test_fp16_posteriors_finite_bounded_never_negative_zero(oracle, kind, rows, iters):
    lam = _inputs(oracle, kind, 11 + rows)
    bits = np.zeros(NCB * 22 * Z, np.int8)
    it = np.zeros(NCB, np.uint8)
    post = np.zeros(NCB * 68 * Z, np.float32)
    oracle.or_ldpc_decode_posterior(Z, NCB, ptr(lam), FP16, iters, rows, 22 * Z, 0, ptr(bits), ptr(it),
                                    ptr(post))
    p = post.reshape(NCB, 68 * Z)[:, : (22 + rows) * Z]
    assert np.isfinite(p).all()
    assert not np.any((p == 0) & np.signbit(p)), "a posterior is -0"
    assert np.abs(p).max() < 32768.0, float(np.abs(p).max())
