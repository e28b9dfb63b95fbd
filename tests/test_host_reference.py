"""CPU checks of the library's bit-exact host references and DMRS-LS contract
validators (the reference's host entry points re_demap_batch::Reference,
layer_demap::ReferenceDemap, descramble_mimo::ReferenceDescramble,
rate_dematch_mimo::ReferenceRateDematch and mimo_dmrs_ls's three
validators): against the reference's own compiled host code (oracle/_ref) at
its fixed profile where that compiles, and against the oracle elsewhere."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import OracleGeom, OrTbPlan, cref, or_tb_plan, ptr
from puschrx import abi, ops
from puschrx import profiles as P

REF_PROFILE = P.REFERENCE_FD16X64


def _ref_plan(oracle_plan_layers, slots):
    p = OrTbPlan()
    p.C, p.Z, p.K, p.Kprime, p.filler, p.N, p.Ncb, p.k0 = 143, 384, 8448, 8448, 0, 25344, 25344, 0
    p.tbs = 143 * 8448
    units = slots * 19152
    p.n_low = 143 - units % 143
    p.E_low = 8 * oracle_plan_layers * (units // 143)
    p.E_high = 8 * oracle_plan_layers * (-(-units // 143))
    return p


def test_re_demap_reference_matches_reference_code(lib, reflib):
    cfg = REF_PROFILE.config()
    g = OracleGeom(cfg)
    nr = 3
    rng = np.random.default_rng(1)
    fre = rng.integers(0, 0x7bff, (nr, 14, g.fft), dtype=np.uint16)
    fim = rng.integers(0, 0x7bff, (nr, 14, g.fft), dtype=np.uint16)
    idx = np.zeros(g.pad, np.uint32)
    assert lib.airan_re_demap_batch_BuildGatherIndex(C.byref(cfg), idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                     g.pad) == abi.OK
    ref_idx = np.zeros(g.pad, np.uint32)
    assert reflib.ref_re_demap_gather_index(ptr(ref_idx), g.pad) == 0
    assert np.array_equal(idx, ref_idx)
    got_r = np.full((nr, 14, g.pad), 0x5555, np.uint16)
    got_i = got_r.copy()
    want_r = np.zeros_like(got_r)
    want_i = np.zeros_like(got_r)
    assert lib.airan_re_demap_batch_Reference(C.byref(cfg), ptr(fre), ptr(fim), nr, ptr(idx), ptr(got_r),
                                              ptr(got_i)) == abi.OK
    assert reflib.ref_re_demap_reference(ptr(fre), ptr(fim), nr, ptr(ref_idx), ptr(want_r), ptr(want_i)) == 0
    assert np.array_equal(got_r, want_r) and np.array_equal(got_i, want_i)
    # Aliased output and an odd byte offset are rejected, as in the reference.
    assert lib.airan_re_demap_batch_Reference(C.byref(cfg), ptr(fre), ptr(fim), nr, ptr(idx), ptr(fre),
                                              ptr(got_i)) == abi.INVALID_ARGUMENT
    bad = idx.copy()
    bad[7] |= 1
    assert lib.airan_re_demap_batch_Reference(C.byref(cfg), ptr(fre), ptr(fim), nr, ptr(bad), ptr(got_r),
                                              ptr(got_i)) == abi.INVALID_ARGUMENT


@pytest.mark.parametrize("prof", [REF_PROFILE, P.FULL_CARRIER_4RX, P.MID_4RX_RANK2], ids=lambda p: p.name)
def test_layer_demap_reference(lib, oracle, reflib, prof):
    cfg = prof.config()
    g = OracleGeom(cfg)
    lay = ops.layout(cfg)
    rng = np.random.default_rng(2)
    llr = rng.integers(-2560, 2561, g.nl * g.qm * g.qam_plane, dtype=np.int16)
    got = np.full(g.qm * g.cw_stride, 77, np.int16)
    assert lib.airan_layer_demap_ReferenceDemap(ptr(llr), C.byref(cfg), C.byref(lay), ptr(got)) == abi.OK
    want = np.zeros_like(got)
    oracle.or_layer_demap(cref(cfg), ptr(llr), ptr(want))
    assert np.array_equal(got, want)
    if prof is REF_PROFILE:
        ref = np.zeros_like(got)
        assert reflib.ref_layer_demap_reference(ptr(llr), cref(cfg), ptr(ref)) == 0
        assert np.array_equal(got, ref)
    wrong = abi.PuschMimoLayout.from_buffer_copy(lay)
    wrong.data_stride += 64
    assert lib.airan_layer_demap_ReferenceDemap(ptr(llr), C.byref(cfg), C.byref(wrong), ptr(got)) == \
        abi.LAYOUT_MISMATCH


@pytest.mark.parametrize("prof,slots", [(REF_PROFILE, 3), (P.FULL_CARRIER_4RX, 2), (P.MID_4RX_RANK2, 1)],
                         ids=lambda v: getattr(v, "name", str(v)))
def test_descramble_reference_continues_across_slots(lib, oracle, prof, slots):
    cfg = prof.config()
    g = OracleGeom(cfg)
    lay = ops.layout(cfg)
    rng = np.random.default_rng(3)
    x = rng.integers(-32768, 32768, slots * g.qm * g.cw_stride, dtype=np.int16)
    got = np.zeros_like(x)
    assert lib.airan_descramble_mimo_ReferenceDescramble(ptr(x), None, C.byref(cfg), C.byref(lay), slots,
                                                         ptr(got)) == abi.OK
    want = np.zeros_like(x)
    oracle.or_descramble(cref(cfg), slots, 0, ptr(x), ptr(want))
    # Codeword tails are zero in both.
    assert np.array_equal(got, want)
    sign = np.zeros_like(x)
    assert lib.airan_descramble_mimo_BuildGoldSign(C.byref(cfg), C.byref(lay), slots,
                                                   sign.ctypes.data_as(C.POINTER(C.c_int16)), sign.size) == abi.OK
    got2 = np.zeros_like(x)
    assert lib.airan_descramble_mimo_ReferenceDescramble(ptr(x), ptr(sign), C.byref(cfg), C.byref(lay), slots,
                                                         ptr(got2)) == abi.OK
    assert np.array_equal(got2, want)
    assert lib.airan_descramble_mimo_ReferenceDescramble(ptr(x), None, C.byref(cfg), C.byref(lay), slots,
                                                         ptr(x)) == abi.INVALID_ARGUMENT


def test_rate_dematch_reference_profile(lib, oracle):
    """The reference's fixed plan (C = 143, code blocks spanning slots)."""
    cfg = REF_PROFILE.config()
    g = OracleGeom(cfg)
    lay = ops.layout(cfg)
    slots = 3
    meta = abi.RateDematchKernelMetadata()
    desc = (abi.RateMatchDescriptor * 143)()
    assert lib.airan_rate_dematch_mimo_BuildCurrentProfile(C.byref(cfg), slots, C.byref(abi.PuschMimoLayout()),
                                                           C.byref(meta), desc, 143) == abi.OK
    rng = np.random.default_rng(4)
    x = rng.integers(-2560, 2561, slots * g.qm * g.cw_stride, dtype=np.int16)
    got = np.zeros(143 * 68 * 384, np.int16)
    assert lib.airan_rate_dematch_mimo_ReferenceRateDematch(ptr(x), C.byref(cfg), C.byref(lay), slots, desc, 143,
                                                            ptr(got)) == abi.OK
    want = np.zeros_like(got)
    oracle.or_rate_dematch(C.byref(_ref_plan(g.nl, slots)), g.nl, g.qm, g.cw_symbols, g.cw_stride, slots, ptr(x),
                           ptr(want))
    assert np.array_equal(got, want)
    desc[5].e += g.qm
    assert lib.airan_rate_dematch_mimo_ReferenceRateDematch(ptr(x), C.byref(cfg), C.byref(lay), slots, desc, 143,
                                                            ptr(got)) == abi.LAYOUT_MISMATCH


def test_rate_dematch_reference_for_tb_with_filler(lib, oracle):
    prof = P.FULL_CARRIER_4RX
    cfg = prof.config()
    g = OracleGeom(cfg)
    lay = ops.layout(cfg)
    tbs = ops.tbs(prof.num_rb, 12 * g.ndata, g.qm, g.nl, prof.code_rate)
    plan = ops.tb_plan(tbs, g.cw_symbols * g.qm, g.qm, g.nl)
    assert plan.filler > 0 and plan.E_high < plan.Ncb
    desc = (abi.RateMatchDescriptor * plan.C)()
    assert lib.airan_rate_dematch_mimo_BuildProfileForTb(C.byref(cfg), 1, C.byref(plan), C.byref(abi.PuschMimoLayout()),
                                                         C.byref(abi.RateDematchKernelMetadata()), desc,
                                                         plan.C) == abi.OK
    rng = np.random.default_rng(5)
    x = rng.integers(-2560, 2561, g.qm * g.cw_stride, dtype=np.int16)
    got = np.zeros(plan.C * 68 * plan.Z, np.int16)
    assert lib.airan_rate_dematch_mimo_ReferenceRateDematchForTb(ptr(x), C.byref(cfg), C.byref(lay), 1,
                                                                 C.byref(plan), desc, plan.C, ptr(got)) == abi.OK
    want = np.zeros_like(got)
    op = or_tb_plan(oracle, plan.tbs, plan.G, g.qm, g.nl)
    oracle.or_rate_dematch(C.byref(op), g.nl, g.qm, g.cw_symbols, g.cw_stride, 1, ptr(x), ptr(want))
    assert np.array_equal(got, want)


def _dmrs_ls_profile(lib, cfg, g):
    count = np.zeros(16, np.uint16)
    sc = np.zeros(4 * g.ndmrs * g.pilot_pad, np.uint16)
    assert lib.airan_mimo_dmrs_ls_BuildCurrentProfile(C.byref(cfg), C.byref(abi.PuschMimoLayout()),
                                                      C.byref(abi.MimoDmrsLsKernelMetadata()),
                                                      count.ctypes.data_as(C.POINTER(C.c_uint16)),
                                                      sc.ctypes.data_as(C.POINTER(C.c_uint16))) == abi.OK
    return count, sc


def _cases(count, sc, nl, nd, pilot_pad):
    yield "profile", count, sc
    c = count.copy()
    c[1] += 1                         # unequal over DMRS symbols
    yield "count", c, sc
    s = sc.copy()
    s[3] += 2                         # a pilot off its comb
    yield "sc", count, s
    s = sc.copy()
    s[int(count[0])] = 9              # a non-zero tail entry
    yield "tail", count, s
    if nl > 2:
        c = count.copy()
        c[2 * nd:3 * nd] = 7          # an unknown observation count
        yield "unknown", c, sc


def test_dmrs_ls_validators_match_reference_code(lib, reflib):
    cfg = REF_PROFILE.config()
    g = OracleGeom(cfg)
    count, sc = _dmrs_ls_profile(lib, cfg, g)
    for rank in (1, 2, 4):
        for name, c, s in _cases(count, sc, rank, g.ndmrs, g.pilot_pad):
            mo = np.zeros(4, np.uint32)
            mr = np.zeros(4, np.uint32)
            a = lib.airan_mimo_dmrs_ls_DescribeObservationModels(None, ptr(c), rank, g.ndmrs, ptr(mo))
            b = reflib.ref_dmrs_ls_describe(ptr(c), rank, g.ndmrs, ptr(mr))
            assert a == b, (rank, name)
            if a == 0:
                assert np.array_equal(mo[:rank], mr[:rank])
            # With the config: the same answers at the reference profile.
            assert lib.airan_mimo_dmrs_ls_DescribeObservationModels(cref(cfg), ptr(c), rank, g.ndmrs, ptr(mo)) == b
            assert lib.airan_mimo_dmrs_ls_ValidateNaturalLmmseContract(None, ptr(c), ptr(s), rank, g.ndmrs) == \
                reflib.ref_dmrs_ls_validate_natural(ptr(c), ptr(s), rank, g.ndmrs), (rank, name)
            assert lib.airan_mimo_dmrs_ls_ValidateLmmse798Compatibility(None, ptr(c), rank, g.ndmrs) == \
                reflib.ref_dmrs_ls_validate_798(ptr(c), rank, g.ndmrs), (rank, name)
    # Wrong DMRS symbol count.
    assert lib.airan_mimo_dmrs_ls_DescribeObservationModels(None, ptr(count), 4, 3, ptr(np.zeros(4, np.uint32))) == \
        reflib.ref_dmrs_ls_describe(ptr(count), 4, 3, ptr(np.zeros(4, np.uint32)))


def test_dmrs_ls_validators_general_geometry(lib):
    """At 273 PRB the comb carries 1638 pilots and an FD-OCC2 pair 819: the
    profile's own outputs satisfy the natural contract; only the rank-1/2
    comb-only layers pass the 798-style check (FD-OCC2 layers return -3)."""
    cfg = P.FULL_CARRIER_4RX.config()
    g = OracleGeom(cfg)
    count, sc = _dmrs_ls_profile(lib, cfg, g)
    models = np.zeros(4, np.uint32)
    assert lib.airan_mimo_dmrs_ls_DescribeObservationModels(cref(cfg), ptr(count), 4, g.ndmrs, ptr(models)) == 0
    assert models.tolist() == [abi.OBS_FD_OCC2] * 4
    assert lib.airan_mimo_dmrs_ls_ValidateNaturalLmmseContract(cref(cfg), ptr(count), ptr(sc), 4, g.ndmrs) == 0
    assert lib.airan_mimo_dmrs_ls_ValidateLmmse798Compatibility(cref(cfg), ptr(count), 4, g.ndmrs) == \
        abi.LAYOUT_MISMATCH
    for name, c, s in list(_cases(count, sc, 4, g.ndmrs, g.pilot_pad))[1:]:
        assert lib.airan_mimo_dmrs_ls_ValidateNaturalLmmseContract(cref(cfg), ptr(c), ptr(s), 4, g.ndmrs) == \
            abi.UNSUPPORTED_PROFILE, name
