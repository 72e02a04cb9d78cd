"""The pseudopotential generator (lira_pseudo_generate, host code) against the reference's pinned generation records
(crates/validate/references.md:2147-2170, compared verbatim by crates/validate/tests/pseudo_gen.rs
generation_record_table_matches_references_md): per element the ladder entry kept (f, local), r_c per channel, Z_v,
E_KB and GSS per projector, r_nlc and r_core, printed to 3-4 decimals there and matched here to their last digit; the
refusals with their reasons, and the T5 value of every refused ladder entry of Be and Al. Runs on the CPU.

The gates the generator restates are G1 and T5. Nitrogen is the one element whose ladder choice they do not fix: the
reference refuses f = 1.4 with the polynomial local through a gate of the KB atom's transfer configurations (T3, T4,
T6 or T7), so the free ladder here keeps that entry, and the record of the reference's entry (f = 1.2, p local) is
checked by pinning it.
"""
import json
import os
import re

import pytest

import lira

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pseudo_generation_records.json")))
H = 0.25


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


@pytest.mark.parametrize("z", sorted(int(z) for z in GOLD["accepted"]))
def test_generation_record_matches_the_pinned_row(z):
    row = GOLD["accepted"][str(z)]
    if z == 7:
        # the free ladder keeps f = 1.4 polynomial (module doc); the reference's entry, pinned, gives its row
        assert lira.generate_pseudo(z, H).record["f"] == 1.4
        sp = lira.generate_pseudo(z, H, lira.F_LADDER.index(row["f"]), lira.PSEUDO_LOCAL_HIGHEST)
    else:
        sp = lira.generate_pseudo(z, H)
    _check_row(sp, row)


@pytest.mark.parametrize("z", GOLD["refused_two_valence_s"])
def test_two_valence_s_states_are_refused(z):
    with pytest.raises(lira.Numerical, match="two valence states of l = 0"):
        lira.generate_pseudo(z, H)


@pytest.mark.parametrize("z", [4, 13])
def test_every_ladder_entry_of_the_t5_refusals(z):
    with pytest.raises(lira.Numerical, match="no ladder entry passes"):
        lira.generate_pseudo(z, H)
    for f, local, l, mha in GOLD["refused_t5_mha"][str(z)]:
        fi = lira.F_LADDER.index(float(f))
        lo = lira.PSEUDO_LOCAL_HIGHEST if local == "highest" else lira.PSEUDO_LOCAL_POLYNOMIAL
        with pytest.raises(lira.Numerical) as e:
            lira.generate_pseudo(z, H, fi, lo)
        m = re.search(r"T5 l = (\d): max delta E_equiv ([0-9.]+) mHa", str(e.value))
        assert m, str(e.value)
        assert int(m.group(1)) == l
        assert round(float(m.group(2)), 2) == pytest.approx(mha, abs=1e-9), (f, local, m.group(2))


def test_oxygen_support_radii():
    # r_cmp, the core Hartree tail's 1e-8 Ha radius (references.md:2298), and the derived supports: r_beta = max r_c
    # (r_nlc < max r_c for O), every table's extent gamma_r R + 8h
    sp = lira.generate_pseudo(8, H)
    rec = sp.record
    assert round(rec["r_cmp"], 4) == pytest.approx(GOLD["o_r_cmp"], abs=1e-9)
    rmax = max(rc for _, rc, _, _ in rec["channels"])
    assert rec["projectors"][0][3] == pytest.approx(rmax, rel=1e-12)
    assert sp.projectors[0].radius == pytest.approx(1.5 * rmax + 8 * H, rel=1e-12)
    assert sp.pc_radius == pytest.approx(1.5 * rec["r_pc_support"] + 8 * H, rel=1e-12)
    assert sp.sigma_lr == pytest.approx(2 * H) and sp.du == pytest.approx(H / 256)


def test_filtered_tables_are_band_limited_and_even():
    # every table is a finite Bessel sum below pi/h times the outer step: even in d, so t(k du) - t(0) grows as k^2
    # near the origin (the ratio of the second to the first difference from 0 is 4), and exactly zero past its radius
    sp = lira.generate_pseudo(8, H)
    for t in [sp.v_sr_u, sp.rho_pc_u] + [p.t_u for p in sp.projectors]:
        assert (t[2] - t[0]) / (t[1] - t[0]) == pytest.approx(4.0, rel=1e-3)
        assert t[-1] == 0.0
    # V_SR(0) = V_loc(0) + 2 Z_v / (sqrt(pi) a) is finite and the partial core is positive at the centre
    assert sp.rho_pc_u[0] > 0


def test_pinning_an_entry_and_bad_arguments():
    sp = lira.generate_pseudo(6, H, 2, lira.PSEUDO_LOCAL_HIGHEST)
    assert sp.record["f"] == 1.2 and sp.record["local"] == "p-local"
    with pytest.raises(lira.Invalid):
        lira.generate_pseudo(8, H, 7, -1)
    with pytest.raises(lira.Numerical, match="outside 1..18"):
        lira.generate_pseudo(19, H)
