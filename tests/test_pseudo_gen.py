"""The pseudopotential generator (lira_pseudo_generate, host code) against the reference's pinned generation records
(crates/validate/references.md:2147-2170, compared verbatim by crates/validate/tests/pseudo_gen.rs
generation_record_table_matches_references_md): per element the ladder entry kept (f, local), r_c per channel, Z_v,
E_KB and GSS per projector, r_nlc and r_core, printed to 3-4 decimals there and matched here to their last digit, and
h_req and h_eb (T8's filtered re-run at every ladder spacing); the refusals with their reasons, and the T5 value of every refused ladder entry of Be and Al. Runs on the CPU; the
generations (each runs every gate of the ladder: seconds of host time) are spread over a process pool.
"""
import json
import multiprocessing
import os
import re
from concurrent.futures import ProcessPoolExecutor

import pytest

import lira

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pseudo_generation_records.json")))
H = 0.25


def _generate(args):
    z, fi, lo = args[:3]
    h = args[3] if len(args) > 3 else H
    try:
        sp = lira.generate_pseudo(z, h, fi, lo)
        return {"record": sp.record, "usable": sp.usable, "refusal": sp.refusal, "radius": [p.radius for p in sp.projectors], "pc_radius": sp.pc_radius,
                "sigma_lr": sp.sigma_lr, "du": sp.du,
                "tables": [sp.v_sr_u[:3].tolist() + [float(sp.v_sr_u[-1])],
                           None if sp.rho_pc_u is None else sp.rho_pc_u[:3].tolist() + [float(sp.rho_pc_u[-1])]]
                + [p.t_u[:3].tolist() + [float(p.t_u[-1])] for p in sp.projectors]}
    except lira.QmError as e:
        return {"error": type(e).__name__, "message": str(e)}


@pytest.fixture(scope="module")
def generated():
    jobs = [(z, -1, -1) for z in range(1, 19)]
    for z in ("4", "13"):
        for f, local, _, _ in GOLD["refused_t5_mha"][z]:
            jobs.append((int(z), lira.F_LADDER.index(float(f)),
                         lira.PSEUDO_LOCAL_HIGHEST if local == "highest" else lira.PSEUDO_LOCAL_POLYNOMIAL))
    jobs.append((6, 2, lira.PSEUDO_LOCAL_HIGHEST))
    jobs.append((8, -1, -1, 0.20))
    # spawned workers: a forked child of a process whose OpenMP runtime already started threads (the oracle's tests)
    # can deadlock in it
    ctx = multiprocessing.get_context("spawn")
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1), mp_context=ctx) as pool:
        return dict(zip(jobs, pool.map(_generate, jobs)))


class _Sp:
    def __init__(self, d):
        self.record = d["record"]


def _check_row(sp, row):
    rec = sp.record
    assert rec["f"] == row["f"]
    assert rec["local"] == row["local"]
    assert rec["z_v"] == row["z_v"]
    got_rc = {"spdf"[l]: rc for l, rc, _, _ in rec["channels"]}
    assert set(got_rc) == set(row["r_c"])
    for k, v in row["r_c"].items():
        assert round(got_rc[k], 3) == pytest.approx(v, abs=1e-9), (k, got_rc[k])
    ekb = {"spdf"[l]: e for l, e, _, _ in rec["projectors"]}
    gss = {"spdf"[l]: g for l, _, g, _ in rec["projectors"]}
    assert set(ekb) == set(row["e_kb"])
    for k, v in row["e_kb"].items():
        assert round(ekb[k], 4) == pytest.approx(v, abs=1e-9), (k, ekb[k])
        assert round(gss[k], 4) == pytest.approx(row["gss"][k], abs=1e-9), (k, gss[k])
    if row["r_nlc"] is None:
        assert rec["r_nlc"] is None
    else:
        assert round(rec["r_nlc"], 3) == pytest.approx(row["r_nlc"], abs=1e-9)
    assert round(rec["r_core"], 3) == pytest.approx(row["r_core"], abs=1e-9)
    # T8: the coarsest ladder spacing the filtered tables pass from (h_req) and its egg-box twin (h_eb)
    assert rec["h_req"] == row["h_req"] and rec["h_eb"] == row["h_eb"], (rec["h_req"], rec["h_eb"])


@pytest.mark.parametrize("z", sorted(int(z) for z in GOLD["accepted"]))
def test_generation_record_matches_the_pinned_row(generated, z):
    out = generated[(z, -1, -1)]
    assert "error" not in out, out
    _check_row(_Sp(out), GOLD["accepted"][str(z)])


@pytest.mark.parametrize("z", GOLD["refused_two_valence_s"])
def test_two_valence_s_states_are_refused(generated, z):
    out = generated[(z, -1, -1)]
    assert out.get("error") == "Numerical" and "two valence states of l = 0" in out["message"]


@pytest.mark.parametrize("z", [4, 13])
def test_every_ladder_entry_of_the_t5_refusals(generated, z):
    out = generated[(z, -1, -1)]
    assert out.get("error") == "Numerical" and "no ladder entry passes" in out["message"]
    for f, local, l, mha in GOLD["refused_t5_mha"][str(z)]:
        fi = lira.F_LADDER.index(float(f))
        lo = lira.PSEUDO_LOCAL_HIGHEST if local == "highest" else lira.PSEUDO_LOCAL_POLYNOMIAL
        msg = generated[(z, fi, lo)].get("message", "")
        m = re.search(r"T5 l = (\d): max delta E_equiv ([0-9.]+) mHa", msg)
        assert m, msg
        assert int(m.group(1)) == l
        assert round(float(m.group(2)), 2) == pytest.approx(mha, abs=1e-9), (f, local, m.group(2))


def test_oxygen_support_radii(generated):
    # r_cmp, the core Hartree tail's 1e-8 Ha radius (references.md:2298), and the derived supports: r_beta = max r_c
    # (r_nlc < max r_c for O), every table's extent gamma_r R + 8h
    out = generated[(8, -1, -1)]
    rec = out["record"]
    assert round(rec["r_cmp"], 4) == pytest.approx(GOLD["o_r_cmp"], abs=1e-9)
    rmax = max(rc for _, rc, _, _ in rec["channels"])
    assert rec["projectors"][0][3] == pytest.approx(rmax, rel=1e-12)
    assert out["radius"][0] == pytest.approx(1.5 * rmax + 8 * H, rel=1e-12)
    assert out["pc_radius"] == pytest.approx(1.5 * rec["r_pc_support"] + 8 * H, rel=1e-12)
    assert out["sigma_lr"] == pytest.approx(2 * H) and out["du"] == pytest.approx(H / 256)


def test_filtered_tables_are_band_limited_and_even(generated):
    # every table is a finite Bessel sum below pi/h times the outer step: even in d, so t(k du) - t(0) grows as k^2
    # near the origin (the ratio of the second to the first difference from 0 is 4), and exactly zero past its radius
    out = generated[(8, -1, -1)]
    for t in out["tables"]:
        assert (t[2] - t[0]) / (t[1] - t[0]) == pytest.approx(4.0, rel=1e-3)
        assert t[3] == 0.0
    # the partial core is positive at the centre
    assert out["tables"][1][0] > 0


def test_pinning_an_entry_and_bad_arguments(generated):
    rec = generated[(6, 2, lira.PSEUDO_LOCAL_HIGHEST)]["record"]
    assert rec["f"] == 1.2 and rec["local"] == "p-local"
    with pytest.raises(lira.Invalid):
        lira.generate_pseudo(8, H, 7, -1)
    with pytest.raises(lira.Numerical, match="outside 1..18"):
        lira.generate_pseudo(19, H)


def test_ionic_check_of_a_table_at_its_spacing(generated):
    # qm::ionic::check (crates/qm/src/ionic.rs:84-102; its test at :256-262): O's h_req is 0.20, so its table filtered
    # for 0.25 is too coarse and the one for 0.20 passes its own filtered re-run; the ungated H is provisional
    o25, o20, h25 = generated[(8, -1, -1)], generated[(8, -1, -1, 0.20)], generated[(1, -1, -1)]
    assert not o25["usable"] and "coarser than the pseudopotential allows" in o25["refusal"]
    assert o20["usable"] and o20["refusal"] == ""
    assert not h25["usable"] and "provisional" in h25["refusal"]
