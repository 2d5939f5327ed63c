"""GPU tests of the serving loop (cbs_server_start / cbs_submit / cbs_wait).

* every request's result is the same alone (max_active = 1) and in a mix of
  grids, methods, precisions, averaging modes and tolerances, and the same as
  a direct cbs_solve of that request;
* a fixed-iteration FP64 Born request matches the oracle to 1e-10;
* admission by HBM footprint: with a budget that holds one request at a time
  every request still completes with the same result; a request that can
  never fit is refused at submit;
* cancel (queued and running), stop, poll and error paths.
"""

import numpy as np
import pytest

from cases import N_BG, RES, oracle_setup, random_field, rel, spec_from

pytestmark = pytest.mark.gpu


def _request(fm, oracle, grid, bc, seed):
    s, eps, init = oracle_setup(oracle, grid, bc, seed)
    src = oracle.PlaneWave(wavelength=500e-9, polarization=[1, 0, 0], k_vector=[0, 0, 1])
    E_inc, _ = oracle.incident_field(s, src, N_BG ** 2)
    S = oracle.source_term(init.potential, E_inc, init.eps_imag)
    return init, spec_from(fm, init), S


def _mix(fm, oracle):
    absorb = oracle.AbsorbingBC(thickness=300e-9, attenuation_thickness=200e-9)
    per = oracle.PeriodicBC()
    out = []
    # (grid, bc, seed, opts)
    for grid, bc, seed, kw in [
        ((16, 16, 12), absorb, 31, dict(method="born", precision="fp64", tol=2e-6, maxiters=3000)),
        ((32, 32, 16), absorb, 32, dict(method="born", precision="fp32_refine", tol=2e-6,
                                         maxiters=5000)),
        ((24, 16, 12), absorb, 33, dict(method="bicgstab", precision="fp64", tol=1e-9,
                                         maxiters=400)),
        ((16, 20, 8), per, 34, dict(method="born", precision="fp64", tol=0.0, maxiters=37)),
        ((20, 16, 12), absorb, 35, dict(method="bicgstab", precision="fp32_refine", tol=1e-7,
                                         maxiters=600, stop_rule="krylov", abstol=0.0)),
        ((12, 12, 10), per, 36, dict(method="born", precision="fp32", tol=0.0, maxiters=25)),
    ]:
        init, spec, S = _request(fm, oracle, grid, bc, seed)
        out.append((init, spec, S, fm.make_opts(h_scale=0.3 - 0.2j, **kw)))
    return out


def _serve(fm, reqs, max_active, policy="lead"):
    srv = fm.CBSServer(max_active=max_active, policy=policy)
    try:
        if max_active == 1:
            res = []
            for init, spec, S, o in reqs:
                t = srv.submit(init.potential, spec, S, o)
                res.append(srv.wait(t))
            return res
        ts = [srv.submit(init.potential, spec, S, o) for init, spec, S, o in reqs]
        return [srv.wait(t) for t in ts]
    finally:
        srv.stop()


@pytest.mark.parametrize("policy", ["lead", "share"])
def test_serve_same_alone_and_in_mix(fm, oracle, policy):
    """Each request's fields are the same bits alone and in a mix, under both
    chunk schedules (one request per chunk, or equal shares in shared launches)."""
    print(f"case serve alone vs mix, policy {policy}", flush=True)
    reqs = _mix(fm, oracle)
    alone = _serve(fm, reqs, 1, policy)
    mixed = _serve(fm, reqs, 8, policy)
    for i, ((E1, H1, s1), (E2, H2, s2)) in enumerate(zip(alone, mixed)):
        print(f"  request {i}: iters {s1.iters} / {s2.iters}, resid {s1.resid:.3e}", flush=True)
        assert s1.iters == s2.iters and s1.converged == s2.converged
        assert s1.refine_steps == s2.refine_steps
        assert np.array_equal(E1, E2), f"request {i}: E differs, rel {rel(E2, E1):.3e}"
        assert np.array_equal(H1, H2), f"request {i}: H differs"
        assert s1.resid == s2.resid
    # the same requests through the direct solve entry point
    for i, ((init, spec, S, o), (E1, H1, s1)) in enumerate(zip(reqs, alone)):
        prob = fm.CBSProblem(init.potential, spec,
                             precision="fp64" if o.precision == 0 else "fp32")
        E, H, st = prob.solve(S, o)
        assert st.iters == s1.iters
        assert np.array_equal(E, E1), f"request {i}: served != cbs_solve, rel {rel(E1, E):.3e}"
        assert np.array_equal(H, H1)
        if o.tol > 0 and o.precision != 1:
            assert s1.converged == 1


def test_serve_fixed_iterations_vs_oracle(fm, oracle):
    print("case serve vs oracle", flush=True)
    bc = oracle.AbsorbingBC(thickness=300e-9, attenuation_thickness=200e-9)
    init, spec, S = _request(fm, oracle, (16, 16, 12), bc, seed=41)
    rhs = oracle.green_conv(S, init.green)
    k = 11
    x_o, st_o = oracle.solve_born(rhs, rhs, init.potential, init.eps_imag, init.green, 0.0, k)
    with fm.CBSServer(max_active=4) as srv:
        t = srv.submit(init.potential, spec, S,
                       fm.make_opts(method="born", tol=0.0, maxiters=k, h_scale=1j))
        E, H, st = srv.wait(t)
    assert st.iters == k and st.converged == 0
    assert rel(E, x_o) <= 1e-10
    assert rel(H, 1j * oracle.curl_conv(x_o, RES)) <= 1e-10
    assert st.voxel_iters == k * int(np.prod(init.potential.shape))


def test_serve_admission_by_footprint(fm, oracle):
    print("case serve admission", flush=True)
    reqs = _mix(fm, oracle)[:3]
    fps = [fm.request_footprint(spec, o) for _, spec, _, o in reqs]
    ref = _serve(fm, reqs, 8)
    # a context whose budget holds the largest request but never two at once
    ctx = fm.Context(0, hbm_budget=int(max(fps) * 1.5))
    try:
        srv = fm.CBSServer(max_active=8, ctx=ctx)
        try:
            ts = [srv.submit(init.potential, spec, S, o) for init, spec, S, o in reqs]
            out = [srv.wait(t) for t in ts]
            for (E1, _, s1), (E2, _, s2) in zip(ref, out):
                assert s1.iters == s2.iters
                assert np.array_equal(E1, E2)
            # a request larger than the whole budget is refused at submit
            big = _request(fm, oracle, (48, 48, 32), oracle.PeriodicBC(), seed=5)
            with pytest.raises(MemoryError):
                srv.submit(big[0].potential, big[1], big[2], fm.make_opts(method="bicgstab"))
        finally:
            srv.stop()
    finally:
        ctx.close()


def test_serve_cancel_stop_poll(fm, oracle):
    print("case serve cancel", flush=True)
    init, spec, S = _request(fm, oracle, (16, 16, 12), oracle.PeriodicBC(), seed=51)
    long_opts = fm.make_opts(method="born", tol=0.0, maxiters=200000)
    short_opts = fm.make_opts(method="born", tol=0.0, maxiters=6)
    srv = fm.CBSServer(max_active=1)
    try:
        a = srv.submit(init.potential, spec, S, long_opts)   # runs
        b = srv.submit(init.potential, spec, S, short_opts)  # queued behind a
        done, _ = srv.poll(b)
        assert not done
        srv.cancel(b)
        srv.cancel(a)
        for t in (a, b):
            with pytest.raises(fm.CBSError) as ei:
                srv.wait(t)
            assert ei.value.code == fm._lib.CBS_ECANCELED
        # the loop keeps serving after cancellations
        c = srv.submit(init.potential, spec, S, short_opts)
        E, _, st = srv.wait(c)
        assert st.iters == 6
        done, st2 = srv.poll(srv.submit(init.potential, spec, S, short_opts))
        with pytest.raises(fm.CBSError) as ei:
            srv.wait(123456789)
        assert ei.value.code == fm._lib.CBS_ENOTFOUND
        with pytest.raises(ValueError):
            srv.submit(init.potential, spec, S, fm.make_opts(method="gmres"))
        d = srv.submit(init.potential, spec, S, long_opts)
    finally:
        srv.stop()
    # stop ends the running request as canceled
    with pytest.raises(fm.CBSError) as ei:
        srv.wait(d)
    assert ei.value.code == fm._lib.CBS_ECANCELED


def _scene_requests(fm, oracle):
    """Scene requests of different grids, boundaries, sources and methods."""
    from cases import scene_bcs, scene_medium, scene_sources
    from frequencymaxwell import solver as S

    out = []
    for grid, kind, many, seed, kw in [
        ((16, 12, 10), "periodic", True, 61, dict(method="born", precision="fp64", tol=0.0,
                                                   maxiters=40)),
        ((12, 12, 10), "absorbing", False, 62, dict(method="born", precision="fp32_refine",
                                                     tol=2e-6, maxiters=3000)),
        ((20, 12, 8), "z_exponential", True, 63, dict(method="bicgstab", precision="fp64",
                                                       tol=1e-8, maxiters=400)),
        ((16, 16, 12), "absorbing", False, 64, dict(method="born", precision="fp64", tol=1e-15,
                                                     maxiters=0)),
    ]:
        bf, bo = scene_bcs(fm, oracle, kind)
        srcf, srco = scene_sources(fm, oracle, many)
        med = scene_medium(fm, grid, seed)
        solver = fm.ConvergentBornSolver(resolution=RES, grid_size=grid, boundary_conditions=bf)
        d, eps, waves, keep = S.scene(solver, srcf, med)
        out.append((d, eps, waves, keep, fm.make_opts(stop_rule="rhs", **kw), (grid, bo, srco, med)))
    return out


def test_serve_scenes_alone_mix_and_direct(fm, oracle):
    """Scene requests (device-side setup and teardown) give the same cropped
    total fields alone, in a mix and through cbs_solve_scene; the fixed-budget
    one matches the oracle's whole solve."""
    print("case serve scenes", flush=True)
    from frequencymaxwell import engine

    reqs = _scene_requests(fm, oracle)
    res = {}
    for active in (1, 8):
        srv = fm.CBSServer(max_active=active)
        try:
            if active == 1:
                res[active] = [srv.wait(srv.submit_scene(d, e, w, o)) for d, e, w, _, o, _ in reqs]
            else:
                ts = [srv.submit_scene(d, e, w, o) for d, e, w, _, o, _ in reqs]
                res[active] = [srv.wait(t) for t in ts]
        finally:
            srv.stop()
    for i, (d, e, w, _, o, meta) in enumerate(reqs):
        (E1, H1, s1), (E2, H2, s2) = res[1][i], res[8][i]
        print(f"  scene {i}: iters {s1.iters}", flush=True)
        assert E1.shape == meta[0] + (3,)
        assert s1.iters == s2.iters
        assert np.array_equal(E1, E2) and np.array_equal(H1, H2), f"scene {i} differs in the mix"
        E3, H3, s3, info = engine.solve_scene(d, e, w, o)
        assert s3.iters == s1.iters
        assert np.array_equal(E3, E1) and np.array_equal(H3, H1), f"scene {i}: served != direct"
        if i == 3:  # the reference budget max(2 Bornmax, 100)
            assert s1.iters == info.maxiters
    grid, bo, srco, med = reqs[0][5]
    s = oracle.Solver(resolution=RES, grid_size=grid, boundary_conditions=bo, tolerance=0.0)
    ref = oracle.solve(s, srco, med.permittivity, N_BG ** 2, method="born", rule="rhs", maxiters=40)
    E1, H1, _ = res[1][0]
    assert rel(E1, ref.E) <= 1e-10
    assert rel(H1, ref.H) <= 1e-10


def test_serve_policy_argument_checks(fm):
    print("case serve policy checks", flush=True)
    with pytest.raises(ValueError):
        fm.CBSServer(max_active=2, policy="fastest")
    srv = fm.CBSServer(max_active=2)
    try:
        with pytest.raises(ValueError):  # CBS_EINVAL
            fm._lib.check(fm._lib.load().cbs_server_policy(srv.ctx.handle, 7, 0.2))
        with pytest.raises(ValueError):
            fm._lib.check(fm._lib.load().cbs_server_policy(srv.ctx.handle, 1, -1.0))
        fm._lib.check(fm._lib.load().cbs_server_policy(srv.ctx.handle, 0, 0.0))
    finally:
        srv.stop()
