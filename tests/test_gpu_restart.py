"""Checkpoint and resume through the C ABI (trt_restart_out / trt_restart_in, the
reference's RESTART_OUT / RESTART_IN): the file holds the context's arrays in the
reference's record order and shapes, a resumed context continues bitwise as the
one that wrote it, a context resumed from the oracle's state follows the oracle,
and bad files are rejected with the context unchanged."""
import numpy as np
import pytest

import oracle.restart as orst

pytestmark = pytest.mark.gpu

TIGHT = dict(use_line_search=0, conv_ho=1e-11, conv_lo=1e-12, conv_gr1=1e-13)
WATCH = ("Temp", "E_avg", "Eg_avg", "HO_E_avg", "MGQD_E_avg", "fg_avg_xx", "fg_avg_xy", "I_crn", "KapE", "Eg_edgV")


def _prob(trt, **kw):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    for k, v in {**TIGHT, **kw}.items():
        setattr(prob, k, v)
    return prob


def _shapes(prob):
    nx, ny, nm, ng = prob.nx, prob.ny, prob.nm, prob.ng
    cg, vg, hg = (nx, ny, ng), (nx + 1, ny, ng), (nx, ny + 1, ng)
    return {"Time": (1,), "Temp": (nx, ny), "I_crn": (2 * nx, 2 * ny, nm, ng), "fg_avg_xx": cg, "fg_avg_yy": cg,
            "fg_edgV_xx": vg, "fg_edgV_xy": vg, "fg_edgH_yy": hg, "fg_edgH_xy": hg,
            "RT_Src": (2 * nx, 2 * ny, nm, ng), "MGQD_Src": cg, "GQD_Src": (nx, ny), "E_avg": (nx, ny), "Eg_avg": cg,
            "Eg_edgV": vg, "Eg_edgH": hg, "Fxg_edgV": vg, "Fyg_edgH": hg, "KapE_Bar": (nx, ny), "KapE": cg,
            "KapR": cg, "Cg_L": (ny, ng), "Cg_B": (nx, ng), "Cg_R": (ny, ng), "Cg_T": (nx, ng)}


def _read(path, prob):
    sh = _shapes(prob)
    return dict(zip(orst.RECORDS, orst.read_restart(path, [sh[n] for n in orst.RECORDS])))


@pytest.mark.parametrize("precision", [0, 1])
def test_restart_out_holds_the_context_state(trt, tmp_path, precision):
    prob = _prob(trt)
    ctx = trt.native.Context(prob, precision=precision)
    for _ in range(2):
        ctx.step()
    f = str(tmp_path / "ctx.rst")
    ctx.restart_out(f)
    rec = _read(f, prob)
    assert rec["Time"][0] == ctx.time == pytest.approx(2 * prob.dt)
    assert trt.native.restart_time_in(f) == ctx.time
    for n in orst.RECORDS[1:]:
        if n == "RT_Src":
            src = ctx.pull(n).reshape(prob.nx, prob.ny, prob.ng, order="F")
            full = np.repeat(np.repeat(src, 2, axis=0), 2, axis=1)[:, :, None, :]
            assert np.array_equal(rec[n], np.broadcast_to(full, rec[n].shape)), n
        else:
            assert np.array_equal(rec[n].ravel(order="F"), ctx.pull(n)), n
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_resume_continues_bitwise(trt, tmp_path, precision):
    """N steps, checkpoint, M more == checkpoint read by a fresh context, M steps."""
    prob = _prob(trt)
    f = str(tmp_path / "a.rst")
    a = trt.native.Context(prob, precision=precision)
    for _ in range(2):
        a.step()
    a.restart_out(f)
    t_ck = a.time
    its_a = [a.step() for _ in range(2)]
    b = trt.native.Context(prob, precision=precision)
    b.restart_in(f)
    assert b.time == t_ck
    its_b = [b.step() for _ in range(2)]
    assert its_a == its_b
    assert a.time == b.time
    for n in WATCH:
        assert np.array_equal(a.pull(n), b.pull(n)), n
    a.close()
    b.close()


def test_resume_from_oracle_state(trt, oracle_mod, tmp_path):
    """The oracle's state after two steps, written as the reference's restart file, read
    by a context: both then take two steps and agree to the time-step tolerance."""
    prob = _prob(trt)
    o = oracle_mod.Oracle(prob, threads=4)
    for _ in range(2):
        o.step()
    f = str(tmp_path / "o.rst")
    orst.write_restart(f, orst.restart_arrays(o, prob.nm))
    ctx = trt.native.Context(prob)
    ctx.restart_in(f)
    for n in ("Temp", "Eg_avg", "I_crn", "Cg_T"):
        assert np.array_equal(ctx.pull(n), np.asarray(o[n]).ravel(order="F")), n
    for s in range(2):
        it_o = o.step()
        it_g = ctx.step()
        for n in ("Temp", "E_avg", "HO_E_avg", "MGQD_E_avg"):
            ref = np.asarray(o[n]).ravel(order="F")
            err = float(np.sqrt(np.sum((ctx.pull(n) - ref) ** 2) / np.sum(ref ** 2)))
            assert err < 1e-8, (s, n, err, it_o, it_g)
    assert ctx.time == pytest.approx(o.time)
    ctx.close()


def test_restart_in_rejects_bad_files(trt, tmp_path):
    prob = _prob(trt)
    ctx = trt.native.Context(prob)
    ctx.step()
    before = {n: ctx.pull(n) for n in ("Temp", "I_crn", "Cg_T")}
    t0 = ctx.time
    # a file of another grid (Temp is the first record that differs), and one of the
    # same grid with another group count (I_crn is)
    for kw, first_bad in ((dict(nx=8, ny=7, ng=3), r"record 2 \(Temp\): holds 448 bytes, expected 504"),
                          (dict(nx=9, ny=7, ng=2), r"record 3 \(I_crn\)")):
        other = trt.native.Context(trt.problem.small(order=4, **kw))
        f_other = str(tmp_path / "other.rst")
        other.restart_out(f_other)
        other.close()
        with pytest.raises(trt.native.TrtError, match=first_bad):
            ctx.restart_in(f_other)
    # RT_Src that is not uniform over the directions
    good = str(tmp_path / "good.rst")
    ctx.restart_out(good)
    rec = _read(good, prob)
    rec["RT_Src"] = rec["RT_Src"].copy(order="F")
    rec["RT_Src"][3, 2, 1, 0] *= 1.5
    f_rt = str(tmp_path / "rt.rst")
    orst.write_restart(f_rt, [rec[n] for n in orst.RECORDS])
    with pytest.raises(trt.native.TrtError, match="RT_Src"):
        ctx.restart_in(f_rt)
    # a truncated file: the last record is cut
    raw = open(good, "rb").read()
    f_cut = str(tmp_path / "cut.rst")
    open(f_cut, "wb").write(raw[:-12])
    with pytest.raises(trt.native.TrtError, match=r"record 25 \(Cg_T\)"):
        ctx.restart_in(f_cut)
    with pytest.raises(trt.native.TrtError, match="could not be opened"):
        ctx.restart_in(str(tmp_path / "missing.rst"))
    assert ctx.time == t0
    for n, v in before.items():
        assert np.array_equal(ctx.pull(n), v), n
    # and the context still steps
    ctx.step()
    ctx.close()


def test_split_records_resume_bitwise(trt, tmp_path, monkeypatch):
    """Records split into many subrecords (as records above 2 GiB are) read back exactly."""
    prob = _prob(trt)
    a = trt.native.Context(prob)
    a.step()
    monkeypatch.setenv("TRT_SUBRECORD_LENGTH", "1000")
    f = str(tmp_path / "split.rst")
    a.restart_out(f)
    monkeypatch.delenv("TRT_SUBRECORD_LENGTH")
    rec = _read(f, prob)
    assert np.array_equal(rec["I_crn"].ravel(order="F"), a.pull("I_crn"))
    b = trt.native.Context(prob)
    b.restart_in(f)
    a.step()
    b.step()
    for n in WATCH:
        assert np.array_equal(a.pull(n), b.pull(n)), n
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_tgen_resume_bitwise(trt, tmp_path, precision):
    """The sweep-only Tgen_QDf step (the served unit of work) resumes bitwise too."""
    prob = _prob(trt)
    f = str(tmp_path / "tgen.rst")
    a = trt.native.Context(prob, precision=precision)
    for _ in range(2):
        a.tgen_step()
    a.restart_out(f)
    for _ in range(2):
        a.tgen_step()
    b = trt.native.Context(prob, precision=precision)
    b.restart_in(f)
    for _ in range(2):
        b.tgen_step()
    assert a.time == b.time
    for n in ("I_crn", "fg_avg_xx", "fg_avg_xy", "fg_edgV_xy", "fg_edgH_yy", "Cg_L", "Cg_T", "HO_Eg_avg", "KapE"):
        assert np.array_equal(a.pull(n), b.pull(n)), n
    a.close()
    b.close()
