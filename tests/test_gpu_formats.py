"""The reference's own integration case run as its driver runs it: the input
directory test/data/config1 (ControlFile.in + three .pnl hulls) read by
formats.problem_from_control, solved on the GPU (plain FP64 path), and written
by formats.write_hams as OAMASS/ODAMPING/OEXFOR files, which are compared with
the case's expected_output files record by record: identical frequency
columns, and values within the bar of the coefficient test on the same case
(5e-5 of the frequency's largest added-mass or damping value, 2e-5 of its
largest excitation)."""
import os

import numpy as np
import pytest

import formats
import hm_bem

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_config1_input_directory_to_hams_files(tmp_path):
    inp = os.path.join(GOLDEN, "config1_input")
    p, cf = formats.problem_from_control(inp)
    props = [formats.wave_properties(v, cf.inft, cf.ouft, cf.h) for v in cf.values]
    c = hm_bem.Context()
    try:
        c.setup(p)
        amas, bdmp, exfc = c.run_frequencies(p.omegas, hm_bem.HM_FP64)
    finally:
        c.close()
    formats.write_hams(str(tmp_path), [r["wk"] for r in props], [r["w1"] for r in props], amas, bdmp, exfc)
    same = total = 0
    for stem in ("OAMASS", "ODAMPING", "OEXFOR"):
        want_all, got_all = [], []
        for md in range(6):
            name = "%s%d.txt" % (stem, md + 1)
            want = [ln.split() for ln in open(os.path.join(GOLDEN, "config1_expected", name)).read().split("\n") if ln.strip()]
            got = [ln.split() for ln in open(os.path.join(str(tmp_path), name)).read().split("\n") if ln.strip()]
            assert len(got) == len(want) == 27, name
            assert [r[:2] for r in got] == [r[:2] for r in want], name
            want_all.append(np.array([[float(t) for t in r[2:]] for r in want]))
            got_all.append(np.array([[float(t) for t in r[2:]] for r in got]))
            same += sum(a == b for r, q in zip(got, want) for a, b in zip(r[2:], q[2:]))
            total += want_all[-1].size
        wv, gv = np.stack(want_all), np.stack(got_all)          # (mode, record, value)
        for f in range(3):                                      # 9 records per frequency
            blk = slice(9 * f, 9 * f + 9)
            scale = np.max(np.abs(wv[:, blk]))                  # the frequency's whole matrix / force set
            bar = (5e-5 + 1e-6 / scale) if stem != "OEXFOR" else 2e-5
            err = np.max(np.abs(gv[:, blk] - wv[:, blk])) / scale
            assert err <= bar, (stem, f, err)
    print("printed values identical to the expected files: %d of %d" % (same, total), flush=True)


def test_generalized_mode_feeds_the_solve(tmp_path):
    """a generalized mode whose normals are body 1's heave normals
    (gen_mod_1_4.txt) radiates and integrates exactly as heave does: its added
    mass and damping rows and columns and its excitation equal heave's"""
    import shutil
    inp = os.path.join(GOLDEN, "config1_input")
    for f in os.listdir(inp):
        shutil.copy(os.path.join(inp, f), str(tmp_path))
    base, _ = formats.problem_from_control(inp)
    nb = int(base.body_start[1])
    with open(str(tmp_path / "gen_mod_1_4.txt"), "w") as fh:
        fh.write("".join("%d %r\n" % (k + 1, float(base.dxyz_p[k, 2])) for k in range(nb)))
    p, _ = formats.problem_from_control(str(tmp_path))
    c = hm_bem.Context()
    try:
        c.setup(p)
        r = c.solve_frequency(0.06, hm_bem.HM_FP64)
    finally:
        c.close()
    for k in ("amas", "bdmp"):
        a = r[k]
        s = np.max(np.abs(a))
        assert np.max(np.abs(a[:, 3] - a[:, 2])) <= 1e-12 * s, k
        assert np.max(np.abs(a[3, :] - a[2, :])) <= 1e-12 * s, k
    f = r["exfc"]
    assert np.max(np.abs(f[3] - f[2])) <= 1e-12 * np.max(np.abs(f))
    assert np.max(np.abs(r["amas"][:, 3] - r["amas"][:, 4])) > 1e-3 * np.max(np.abs(r["amas"]))


def test_example5_irr_input_directory_to_hams_files(tmp_path):
    """the reference's single-body example with irregular-frequency removal
    (examples/example5_single_body: 234 hull + 383 interior waterplane panels,
    h = 70 m, 10 frequencies, 36 headings) run as its driver runs it
    (formats.run_input_directory: read, solve on the GPU record by record,
    write), against its Hams_format files: identical frequency columns and
    values within 1e-4 of the frequency's largest added-mass, damping or
    excitation value. The files were written by the reference at GRN_KIND = 4
    (its Green-function tables stored in FP32) and the IRR normal equations
    C = A^T A square the system's conditioning: the FP64 oracle itself sits
    2.2e-6 .. 4.4e-5 from them at these ten frequencies
    (tests/test_oracle.py::test_example5_irr_expected_outputs)"""
    inp = os.path.join(GOLDEN, "example5_input")
    recs, amas, bdmp, exfc = formats.run_input_directory(inp, str(tmp_path))
    assert len(recs) == 10 and amas.shape == (10, 6, 6) and exfc.shape == (10, 6, 36)
    for stem in ("OAMASS", "ODAMPING", "OEXFOR"):
        per = 1 if stem != "OEXFOR" else 36
        want_all, got_all = [], []
        for md in range(6):
            name = "%s%d.txt" % (stem, md + 1)
            want = [ln.split() for ln in open(os.path.join(GOLDEN, "example5_expected", name)).read().split("\n")
                    if ln.strip()]
            got = [ln.split() for ln in open(os.path.join(str(tmp_path), "Hams_format", name)).read().split("\n")
                   if ln.strip()]
            assert len(got) == len(want) == 10 * per, name
            assert [r[:2] for r in got] == [r[:2] for r in want], name
            want_all.append(np.array([[float(t) for t in r[2:]] for r in want]))
            got_all.append(np.array([[float(t) for t in r[2:]] for r in got]))
        wv, gv = np.stack(want_all), np.stack(got_all)
        errs = []
        for f in range(10):
            blk = slice(per * f, per * f + per)
            scale = np.max(np.abs(wv[:, blk]))
            errs.append(float(np.max(np.abs(gv[:, blk] - wv[:, blk])) / scale))
        print(stem, " ".join("%.1e" % e for e in errs), flush=True)
        assert max(errs) <= 1e-4, (stem, errs)
    assert os.path.exists(os.path.join(str(tmp_path), "Wamit_format", "AmssDamp.1"))
