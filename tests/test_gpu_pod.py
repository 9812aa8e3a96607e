"""POD reconstruction on the GPU (trt_pod_*) against the oracle's restatement, and a ROM
step driven by it: bases built from the oracle's own snapshots, fields reconstructed into
the *_next arrays on the device, trt_step_given_qdf against the oracle's same step."""
import numpy as np
import pytest

import oracle.rom as rom

pytestmark = pytest.mark.gpu

QDF = ("fg_avg_xx", "fg_avg_yy", "fg_edgV_xx", "fg_edgV_xy", "fg_edgH_yy", "fg_edgH_xy", "Cg_L", "Cg_B", "Cg_R",
       "Cg_T")


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name,gsum,rank", [("fg_avg_xx_next", 0, None), ("fg_edgV_xy_next", 0, 3),
                                            ("Cg_L_next", 1, 4), ("fg_edgH_yy_next", 1, None)])
def test_pod_reconstruct_matches_oracle(trt, name, gsum, rank):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    ctx = trt.native.Context(prob)
    n = ctx.size(name)
    nt = 5
    rng = np.random.default_rng(n + gsum)
    X = rng.uniform(0.2, 0.5, (n, nt))
    C, S, U, V, rk = rom.basis_from_snapshots(X, prob.ng, gsum, rank)
    pod = trt.native.PodBasis(ctx, name, C, S, U, V, rk, nt, gsum)
    for t in (1, 3, 5):
        pod.reconstruct(t)
        ref = rom.pod_reconstruct(C, S, U, V, rk, n, prob.ng, nt, t, gsum)
        assert _rel(ctx.pull(name), ref) < 1e-14, (t, _rel(ctx.pull(name), ref))
        if rank is None:
            assert _rel(ctx.pull(name), X[:, t - 1]) < 1e-12
    pod.reconstruct(2, 3, 0.27, 0.2, 0.3)
    f2 = rom.pod_reconstruct(C, S, U, V, rk, n, prob.ng, nt, 2, gsum)
    f3 = rom.pod_reconstruct(C, S, U, V, rk, n, prob.ng, nt, 3, gsum)
    assert _rel(ctx.pull(name), rom.linear_interpolation(f2, f3, 0.27, 0.2, 0.3)) < 1e-14
    with pytest.raises(trt.native.TrtError, match="out of range"):
        pod.reconstruct(6)
    pod.close()
    with pytest.raises(trt.native.TrtError, match="no array"):
        trt.native.PodBasis(ctx, "fg_avg_zz_next", C, S, U, V, rk, nt, gsum)
    ctx.close()


def test_rom_step_from_pod_bases(trt, oracle_mod):
    """Snapshots of three full oracle steps -> per-group bases (full rank) -> the fields of
    the second snapshot time reconstructed on the GPU -> one ROM step, GPU against oracle."""
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    snap = oracle_mod.Oracle(prob, threads=4)
    X = {n: [] for n in QDF}
    for _ in range(3):
        snap.step()
        for n in QDF:
            X[n].append(np.asarray(snap[n]).ravel(order="F").copy())
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob)
    o.step()
    ctx.step()
    pods = []
    for n in QDF:
        Xn = np.stack(X[n], axis=1)
        C, S, U, V, rk = rom.basis_from_snapshots(Xn, prob.ng, 0)
        pod = trt.native.PodBasis(ctx, n + "_next", C, S, U, V, rk, 3, 0)
        pod.reconstruct(2)
        pods.append(pod)
        ref = rom.pod_reconstruct(C, S, U, V, rk, Xn.shape[0], prob.ng, 3, 2, 0)
        o[n + "_next"][...] = ref.reshape(np.asarray(o[n]).shape, order="F")
        assert _rel(ctx.pull(n + "_next"), ref) < 1e-14
    it_o = o.step_given_qdf()
    it_g = ctx.step_given_qdf()
    assert it_g[3] == 0 and it_g[0] == 1
    for n in ("Temp", "E_avg", "MGQD_E_avg", "Eg_avg"):
        ref = np.asarray(o[n]).ravel(order="F")
        err = float(np.sqrt(np.sum((ctx.pull(n) - ref) ** 2) / np.sum(ref ** 2)))
        assert err < 1e-8, (n, err, it_o, it_g)
    for p in pods:
        p.close()
    ctx.close()


@pytest.mark.parametrize("name,gsum", [("fg_edgV_xy_next", 0), ("Cg_L_next", 1), ("fg_avg_yy_next", 0)])
def test_dmd_reconstruct_matches_oracle(trt, name, gsum):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    ctx = trt.native.Context(prob)
    n = ctx.size(name)
    ng = 1 if gsum else prob.ng
    glen = n // ng
    rng = np.random.default_rng(n + 7 * gsum)
    rank = rng.integers(1, 5, ng)
    tot = int(rank.sum())
    L = rng.uniform(-2.0, 0.5, tot) + 1j * rng.uniform(-3.0, 3.0, tot)
    B = rng.standard_normal(tot) + 1j * rng.standard_normal(tot)
    W = rng.standard_normal(glen * tot) + 1j * rng.standard_normal(glen * tot)
    C = rng.uniform(0.2, 0.4, n)
    dmd = trt.native.DmdBasis(ctx, name, C, L, B, W, rank, gsum)
    for t in (0.0, 0.013, 0.4):
        dmd.reconstruct(t)
        ref = rom.dmd_reconstruct(C, L, B, W, rank, n, prob.ng, t, gsum)
        assert _rel(ctx.pull(name), ref) < 1e-13, (t, _rel(ctx.pull(name), ref))
    dmd.close()
    ctx.close()


def test_bcg_reconstruction_splits_into_the_four_sides(trt):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    ctx = trt.native.Context(prob)
    n = ctx.size("BCg_next")
    assert n == 2 * (prob.nx + prob.ny) * prob.ng
    rng = np.random.default_rng(4)
    X = rng.uniform(0.3, 0.6, (n, 4))
    C, S, U, V, rk = rom.basis_from_snapshots(X, prob.ng, 0)
    pod = trt.native.PodBasis(ctx, "BCg_next", C, S, U, V, rk, 4, 0)
    pod.reconstruct(3)
    bc = rom.pod_reconstruct(C, S, U, V, rk, n, prob.ng, 4, 3, 0)
    assert _rel(ctx.pull("BCg_next"), bc) < 1e-14
    # the split is a copy: bitwise against the split of the device's own vector
    for side, ref in zip(("Cg_L_next", "Cg_B_next", "Cg_R_next", "Cg_T_next"),
                         rom.split_bcg(ctx.pull("BCg_next"), prob.nx, prob.ny, prob.ng)):
        assert np.array_equal(ctx.pull(side), ref.ravel(order="F")), side
    pod.close()
    ctx.close()


def _bcg_vector(L, B, R, T):
    """The reference's C_g boundary vector (POD_ROUTINES.f08:279-352): per group the
    left side (N_y values), the bottom (N_x), the right (N_y), the top (N_x)."""
    ng = L.shape[1]
    return np.concatenate([np.concatenate([L[:, g], B[:, g], R[:, g], T[:, g]]) for g in range(ng)])


def test_bcg_push_splits_in_the_reference_order(trt):
    """A hand-coded vector: each value names its group, side and index, so a wrong side,
    a wrong stride or a transposed group block shows as a wrong code."""
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    nx, ny, ng = prob.nx, prob.ny, prob.ng
    ctx = trt.native.Context(prob)
    code = []
    for g in range(ng):
        for side, n in ((1, ny), (2, nx), (3, ny), (4, nx)):
            code += [g * 1e6 + side * 1e4 + i for i in range(n)]
    ctx.push("BCg_next", np.array(code))
    for side, name, n in ((1, "Cg_L_next", ny), (2, "Cg_B_next", nx), (3, "Cg_R_next", ny), (4, "Cg_T_next", nx)):
        got = ctx.pull(name).reshape((n, ng), order="F")
        want = np.array([[g * 1e6 + side * 1e4 + i for g in range(ng)] for i in range(n)])
        assert np.array_equal(got, want), name
    ctx.close()


@pytest.mark.parametrize("direc_diff", [0, 1])
def test_bcg_push_feeds_the_rom_step(trt, oracle_mod, direc_diff):
    """C_g handed over as the reference's boundary vector (a push of BCg_next), the
    factors pushed one by one; the ROM step against the oracle's step on the same values,
    with Direc_Diff = 1 zeroing the reconstructed off-diagonal face factors."""
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    prob.use_line_search = 0
    prob.conv_ho, prob.conv_lo, prob.conv_gr1 = 1e-11, 1e-12, 1e-13
    o = oracle_mod.Oracle(prob, threads=4)
    ctx = trt.native.Context(prob)
    o.step()
    ctx.step()
    for n in QDF[:6]:
        a = np.asarray(o[n]) * (1.0 + 1e-3 * np.cos(np.arange(np.asarray(o[n]).size) * 0.41).reshape(
            np.asarray(o[n]).shape, order="F"))
        if direc_diff and n in ("fg_edgV_xy", "fg_edgH_xy"):
            o[n + "_next"][...] = 0.0
        else:
            o[n + "_next"][...] = a
        ctx.push(n + "_next", a)
    sides = []
    for n in ("Cg_L", "Cg_B", "Cg_R", "Cg_T"):
        a = np.asarray(o[n]) * 0.97
        o[n + "_next"][...] = a
        sides.append(a)
    ctx.push("BCg_next", _bcg_vector(*sides))
    it_o = o.step_given_qdf()
    it_g = ctx.step_given_qdf(direc_diff=direc_diff)
    for n in QDF:
        assert np.array_equal(ctx.pull(n), np.asarray(o[n]).ravel(order="F")), n
    for n in ("Temp", "E_avg", "MGQD_E_avg", "Eg_avg"):
        ref = np.asarray(o[n]).ravel(order="F")
        err = float(np.sqrt(np.sum((ctx.pull(n) - ref) ** 2) / np.sum(ref ** 2)))
        assert err < 1e-8, (n, err, it_o, it_g)
    ctx.close()


def test_basis_lifetime_and_targets(trt):
    prob = trt.problem.small(nx=9, ny=7, order=4, ng=3)
    ctx = trt.native.Context(prob)
    n = ctx.size("fg_avg_xx_next")
    rng = np.random.default_rng(2)
    C, S, U, V, rk = rom.basis_from_snapshots(rng.uniform(0.2, 0.5, (n, 3)), prob.ng, 0)
    # the intensities are no context array a basis may write
    for name in ("I_crn", "I_crn_old"):
        m = ctx.size(name)
        with pytest.raises(trt.native.TrtError, match="no array"):
            trt.native.PodBasis(ctx, name, np.zeros(m), S, np.zeros(m * 2), V, [2, 2, 2], 3, 0)
    pod = trt.native.PodBasis(ctx, "fg_avg_xx_next", C, S, U, V, rk, 3, 0)
    pod.reconstruct(1)
    ctx.close()                       # frees the open basis with the context
    pod.close()                       # and the wrapper then only drops its handle
    assert pod.h is None
