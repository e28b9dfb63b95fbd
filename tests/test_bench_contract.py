"""bench.py's contract where no GPU run is involved: it refuses more than one
GPU (exit 2, saying why), fails loudly without a GPU (exit 1), and the
algorithmic bytes it states the decoder's roofline on are the workload's
codeword LLRs in (int16, Qm bits per layer-RE) plus the hard bits out
(int8 per bit): 2.52 + 1.09 MB per slot, 230.8 MB per 64-slot launch."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]


def _run(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=ROOT)


def test_refuses_more_than_one_gpu():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.skipif(torch.cuda.is_available(), reason="the behaviour without a GPU")
def test_fails_loudly_without_a_gpu():
    r = _run("--steps", "1", "--warmup", "0", "--no-cpu")
    assert r.returncode == 1
    assert "needs a GPU" in r.stderr


def test_decoder_bytes_of_the_workload():
    import bench
    from puschrx import ops
    from puschrx import profiles as P

    prof = P.FULL_CARRIER_4RX
    g = ops.Geometry(prof.config())
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    plan = ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)
    assert (plan.C, plan.K, plan.Z) == (129, 8448, 384)
    b = bench.stage_bytes(g, plan, 64, True)
    llr_per_slot = 8 * 4 * 12 * 3276 * 2   # 256QAM, 4 layers, 12 data symbols x 3276 subcarriers, int16
    bits_per_slot = 129 * 8448              # C x K hard bits, int8
    assert b["ldpc_decode"] == 64 * (llr_per_slot + bits_per_slot)
    assert round(b["ldpc_decode"] / 1e6, 1) == 230.8
    assert b["rate_dematch"] == 0          # the fused decoder reads the codeword LLRs itself
    assert b["crc_check"] == 64 * bits_per_slot
