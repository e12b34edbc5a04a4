"""Continuous-batching request loop (lbm_serve) on the GPU.

A request's result must not depend on what else is resident: its cavity's
arithmetic is the single-cavity code and its check points depend only on its
own step count, so alone and in a mix the step count, Nu and status are
identical (FP64: bit for bit).
"""
import numpy as np
import pytest

from lbm_ddf_amd import bytes_required, make_params, serve

pytestmark = pytest.mark.gpu


def req(lx, ly, lz, ra, seed, arrival=0.0, max_steps=600, nu_tol=1e-12, window=100, **kw):
    d = dict(lx=lx, ly=ly, lz=lz, rayleigh=ra, seed=seed, arrival_s=arrival, max_steps=max_steps, nu_tol=nu_tol,
             window=window)
    d.update(kw)
    return d


A = req(20, 16, 12, 1e5, 1, amp=1e-2)
B = req(16, 24, 16, 3e5, 2, amp=1e-2)
C = req(24, 12, 20, 1e6, 3, amp=1e-2, max_steps=900)
D = req(12, 12, 12, 5e4, 4, amp=1e-2, max_steps=300)


def same(r1, r2):
    for k in ("status", "steps", "nu_bulk", "nu_top", "nu_bottom", "kinetic_energy"):
        assert r1[k] == r2[k], (k, r1[k], r2[k])
    assert (np.isnan(r1["nu_running"]) and np.isnan(r2["nu_running"])) or r1["nu_running"] == r2["nu_running"]


@pytest.mark.parametrize("precision", ["fp64", "fp16"])
def test_result_same_alone_and_in_mix(precision):
    alone, _ = serve([A], precision=precision, chunk_steps=50)
    assert alone[0]["status"] == "budget" and alone[0]["steps"] == 600
    mix = [dict(B, arrival_s=0.0), dict(A, arrival_s=0.05), dict(C, arrival_s=0.0), dict(D, arrival_s=0.1)]
    res, st = serve(mix, precision=precision, chunk_steps=50)
    same(alone[0], res[1])
    assert st["completed"] == 4 and st["peak_resident"] >= 2
    for r in res:
        assert r["status"] == "budget"
        assert r["finish_s"] >= r["admit_s"] >= r["arrival_s"]


def test_converged_result_same_alone_and_in_mix():
    a = dict(A, nu_tol=0.5, max_steps=5000)
    alone, _ = serve([a], chunk_steps=50)
    assert alone[0]["status"] == "converged" and alone[0]["steps"] % 100 == 0
    res, _ = serve([B, a, C], chunk_steps=50)
    same(alone[0], res[1])


def test_admission_by_hbm_budget():
    need = bytes_required(make_params(16, 16, 16))
    reqs = [req(16, 16, 16, 1e5, s, max_steps=200) for s in range(4)]
    res, st = serve(reqs, chunk_steps=50, hbm_budget=2 * need + need // 2)
    assert st["peak_resident"] == 2
    assert st["completed"] == 4
    admits = sorted(r["admit_s"] for r in res)
    assert admits[2] >= min(r["finish_s"] for r in res)  # third admitted only after a retirement


def test_resident_cells_cap():
    reqs = [req(16, 16, 16, 1e5, s, max_steps=200) for s in range(4)]
    res, st = serve(reqs, chunk_steps=50, max_resident_cells=2 * 16 ** 3)
    assert st["peak_resident"] == 2 and st["completed"] == 4
    alone, _ = serve([reqs[3]], chunk_steps=50)
    same(alone[0], res[3])


def test_reject_divergence_and_time_limit():
    big = req(64, 64, 64, 1e5, 9, max_steps=100)
    blowup = req(12, 12, 12, 1e5, 10, amp=1e4, max_steps=400)  # absurd perturbation: supersonic, then non-finite
    ok = req(12, 12, 12, 1e5, 11, max_steps=200)
    need_small = bytes_required(make_params(12, 12, 12))
    res, st = serve([big, blowup, ok], chunk_steps=50, hbm_budget=3 * need_small)
    assert res[0]["status"] == "rejected"
    assert res[1]["status"] == "diverged"
    assert res[2]["status"] == "budget"
    alone, _ = serve([ok], chunk_steps=50)
    same(alone[0], res[2])
    # a time limit shorter than the work: admitted ones end unfinished, late ones pending
    slow = [req(24, 24, 24, 1e5, 20, max_steps=10 ** 9), req(24, 24, 24, 1e5, 21, arrival=30.0)]
    res, st = serve(slow, chunk_steps=50, time_limit_s=0.5)
    assert res[0]["status"] == "unfinished" and res[0]["steps"] > 0
    assert res[1]["status"] == "pending"
    assert st["unfinished"] == 1 and st["pending"] == 1


def test_fp16_request_outside_the_codec_range_is_rejected_not_run():
    """FP16 storage holds wall temperatures in [-1, 1]: a request outside it
    is rejected at admission (the create-time check), the others run."""
    hot = dict(A, t_hot=1.5, t_cold=0.0)
    res, st = serve([hot, D], precision="fp16", chunk_steps=50)
    assert res[0]["status"] == "rejected"
    assert res[1]["status"] == "budget" and res[1]["steps"] == 300
    assert st["rejected"] == 1 and st["completed"] == 1


def test_request_beyond_the_member_size_is_rejected_not_run():
    """An ensemble member has fewer than 2^31 cells.  A larger in-place FP16
    request fits a 250 GB budget (about 232 GB), so it is rejected at
    admission, before anything is allocated, and the other request runs."""
    big = req(1291, 1291, 1289, 1e5, 9)  # 2,148,351,809 cells
    res, st = serve([big, D], precision="fp16", chunk_steps=50, pattern="aa")
    assert res[0]["status"] == "rejected"
    assert res[1]["status"] == "budget" and res[1]["steps"] == 300
    assert st["rejected"] == 1 and st["completed"] == 1


def test_in_place_serving_gives_the_same_results():
    """FP64 cavities kept in place (one population buffer) give every request
    exactly the A/B result."""
    mix = [dict(B, arrival_s=0.0), dict(A, arrival_s=0.05), dict(C, arrival_s=0.0), dict(D, arrival_s=0.1)]
    ab, _ = serve(mix, chunk_steps=50)
    aa, _ = serve(mix, chunk_steps=50, pattern="aa")
    for r1, r2 in zip(ab, aa):
        same(r1, r2)


def test_in_place_fp16_result_same_alone_and_in_mix():
    """16-bit in place rounds g against other rows than A/B does, so its
    results are its own -- but they do not depend on the mix."""
    alone, _ = serve([A], precision="fp16", chunk_steps=50, pattern="aa")
    mix = [dict(B, arrival_s=0.0), dict(A, arrival_s=0.05), dict(C, arrival_s=0.0), dict(D, arrival_s=0.1)]
    res, st = serve(mix, precision="fp16", chunk_steps=50, pattern="aa")
    same(alone[0], res[1])
    assert st["completed"] == 4


def test_in_place_serving_admits_more_under_one_budget():
    need_ab = bytes_required(make_params(24, 24, 24))
    need_aa = bytes_required(make_params(24, 24, 24, pattern="aa"))
    assert need_aa < 0.6 * need_ab
    reqs = [req(24, 24, 24, 1e5, s, max_steps=200) for s in range(6)]
    budget = 2 * need_ab + need_ab // 2  # two A/B cavities fit, four in place
    res_ab, st_ab = serve(reqs, chunk_steps=50, hbm_budget=budget)
    res_aa, st_aa = serve(reqs, chunk_steps=50, hbm_budget=budget, pattern="aa")
    assert st_ab["peak_resident"] == 2
    assert st_aa["peak_resident"] >= 4
    assert st_aa["completed"] == st_ab["completed"] == 6
    for r1, r2 in zip(res_ab, res_aa):
        same(r1, r2)


def test_request_results_equal_oracle():
    """A served request's diagnostics at its last check equal the oracle's
    after the same number of steps (FP64; the reductions sum in another
    order, hence 1e-11)."""
    import oracle

    res, _ = serve([B, dict(A, arrival_s=0.02), D], chunk_steps=50)
    for q, r in ((A, res[1]), (D, res[2])):
        assert r["status"] == "budget"
        o = oracle.OracleState(q["lx"], q["ly"], q["lz"], rayleigh=q["rayleigh"], prandtl=0.71, t_hot=0.5,
                               t_cold=-0.5)
        o.init_conductive(q["seed"], q["amp"])
        o.run(r["steps"])
        od = o.diag()
        for k in ("nu_bulk", "nu_top", "nu_bottom", "kinetic_energy"):
            assert np.isclose(r[k], od[k], rtol=1e-11, atol=1e-14), (q["seed"], k, r[k], od[k])
