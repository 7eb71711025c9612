"""GPU: the adjoint of the PBF step (adjoint.hip) against the oracle's FP64
adjoint, the checkpointed rollout, density and the torch.autograd harness.

Kernel-level parity: the oracle's VJP (orc_step_vjp_d) runs on the GPU's own
tape (its iterates p^k, permutation and cell offsets, upcast to FP64), so the
two differ only by the GPU's FP32 arithmetic; the chaotic forward map does not
enter.  Gradients are compared as ||g_gpu - g_ref|| / ||g_ref|| (stated per
test) and per direction along 32 seeded unit directions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def _engine(mtlp, p, x, v):
    e = mtlp.Engine(0, p)
    e.reset(len(x), mtlp.capi.SCENE_EMPTY)
    e.upload_state(x, v)
    return e


def _cpu_tape(mtlp, e, step, iterations):
    C = mtlp.capi
    return dict(sorted_index=e.tape_read(C.TAPE_SORTED_INDEX, step),
                cell_start=e.tape_read(C.TAPE_CELL_START, step),
                p=np.ascontiguousarray(np.stack([e.tape_read(C.TAPE_ITERATE, step, k).astype(np.float64)
                                                 for k in range(iterations + 1)])),
                x_in=e.tape_read(C.TAPE_POSITIONS, step).astype(np.float64))


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _scenes(mtlp):
    """(name, params, x, v): the benchmark cube (s_corr k 0.1), the same
    cube with the reference's constants, and a dam column on the floor."""
    out = []
    p = mtlp.scene_params(3, 4096)
    x, v = mtlp.generate_scene(3, 4096, 5, p)
    out.append(("cube_bench", p, x, v))
    q = p.copy()
    q.rest_density, q.epsilon, q.scorr_k = 8000.0, 800.0, 1e-4
    out.append(("cube_reference_constants", q, x, v))
    d = mtlp.scene_params(4, 4096)
    d.iterations = 3
    xd, vd = mtlp.generate_scene(4, 4096, 5, d)
    out.append(("dam_walls", d, xd, vd))
    return out


def _warm(mtlp, p, x, v, steps):
    """State after `steps` plain steps (so walls and motion are engaged)."""
    with _engine(mtlp, p, x, v) as e:
        for _ in range(steps):
            e.step(DT)
        return e.positions_buffer().read(), e.velocities_buffer().read()


def test_taped_forward_is_the_step(mtlp):
    """The taped kernels compute bit for bit what mtlp_step computes."""
    import torch  # noqa: F401
    C = mtlp.capi
    for name, p, x, v in _scenes(mtlp):
        with _engine(mtlp, p, x, v) as e:
            e.tape_record(DT, 3)
            taped = [(e.tape_read(C.TAPE_POSITIONS, t), e.tape_read(C.TAPE_VELOCITIES, t)) for t in range(4)]
            sidx0 = e.tape_read(C.TAPE_SORTED_INDEX, 0)
            cs0 = e.tape_read(C.TAPE_CELL_START, 0)
            np.testing.assert_array_equal(taped[0][0].view(np.uint32), x.view(np.uint32))
            for t in range(3):
                e.step(DT)
                if t == 0:
                    np.testing.assert_array_equal(e.sorted_index_buffer().read(), sidx0, err_msg=name)
                    np.testing.assert_array_equal(e.cell_start_buffer().read(), cs0, err_msg=name)
                xs, vs = e.positions_buffer().read(), e.velocities_buffer().read()
                np.testing.assert_array_equal(xs.view(np.uint32), taped[t + 1][0].view(np.uint32), err_msg=(name, t))
                np.testing.assert_array_equal(vs.view(np.uint32), taped[t + 1][1].view(np.uint32), err_msg=(name, t))
            assert e.error_flags() == 0


def test_rollout_leaves_the_engine_as_steps_do(mtlp):
    """mtlp_rollout (its last checkpoint segment runs taped) leaves every
    engine buffer bit-identical to the same number of mtlp_step calls."""
    p = mtlp.scene_params(4, 4096)
    p.iterations = 3
    x, v = mtlp.generate_scene(4, 4096, 3, p)
    bufs = ("positions", "velocities", "prev_positions", "cell_hashes", "cell_start", "sorted_index",
            "neighbor_counts", "lambdas", "sorted_positions")
    for steps, every in ((5, 2), (6, 3), (1, 4)):
        with _engine(mtlp, p, x, v) as a, _engine(mtlp, p, x, v) as b:
            for _ in range(steps):
                a.step(DT)
            b.rollout(DT, steps, every)
            for name in bufs:
                ra = getattr(a, name + "_buffer")().read()
                rb = getattr(b, name + "_buffer")().read()
                np.testing.assert_array_equal(ra.view(np.uint32), rb.view(np.uint32), err_msg=(steps, every, name))
            assert a.error_flags() == 0 and b.error_flags() == 0


def _step_vjp_check(mtlp, orc, p, x, v, label):
    """One step: GPU (xbar_in, vbar_in, gbar) vs the FP64 oracle adjoint on
    the GPU tape.  Tolerance: 1e-4 relative (norm) per output, 1e-3 along
    each of 32 seeded unit directions relative to the gradient's norm."""
    import torch
    n = len(x)
    rng = np.random.default_rng(11)
    xbar = np.zeros((n, 4), np.float32)
    vbar = np.zeros((n, 4), np.float32)
    xbar[:, :3] = rng.standard_normal((n, 3))
    vbar[:, :3] = rng.standard_normal((n, 3))
    dev = torch.device("cuda")
    with _engine(mtlp, p, x, v) as e:
        s = torch.cuda.current_stream()
        e.rollout(DT, 1, 1, s)
        xb_t, vb_t = torch.from_numpy(xbar).to(dev), torch.from_numpy(vbar).to(dev)
        xo, vo, go = (torch.empty((n, 4), device=dev), torch.empty((n, 4), device=dev),
                      torch.empty(3, device=dev))
        e.adjoint_rollout(DT, 1, 1, xb_t, vb_t, xo, vo, go, s)
        torch.cuda.synchronize()
        gx, gv, gg = xo.cpu().numpy(), vo.cpu().numpy(), go.cpu().numpy()
        e.upload_state(x, v)
        e.tape_record(DT, 1)
        tape = _cpu_tape(mtlp, e, 0, p.iterations)
        assert e.error_flags() == 0
    op = orc.params(p)
    rx, rv, rg = orc.step_vjp_d(tape, op, DT, xbar.astype(np.float64), vbar.astype(np.float64),
                                orc.adjoint_sums(xbar, vbar))
    errs = dict(x=_rel(gx, rx), v=_rel(gv, rv), g=_rel(gg, rg))
    print(label, errs, "|xbar_in|", np.linalg.norm(rx))
    assert np.all(gx[:, 3] == 0) and np.all(gv[:, 3] == 0)
    assert errs["x"] < 1e-4 and errs["v"] < 1e-4 and errs["g"] < 1e-4, errs
    for _ in range(32):
        d = rng.standard_normal((n, 4))
        d[:, 3] = 0
        d /= np.linalg.norm(d)
        assert abs(((gv - rv) * d).sum()) < 1e-3 * np.linalg.norm(rv)
        assert abs(((gx - rx) * d).sum()) < 1e-3 * np.linalg.norm(rx)


@pytest.mark.parametrize("scene", ["cube_bench", "cube_reference_constants", "dam_walls"])
def test_step_vjp_matches_fp64_adjoint(mtlp, orc, scene):
    """One step of a warmed scene against the FP64 adjoint (_step_vjp_check)."""
    name, p, x, v = next(s for s in _scenes(mtlp) if s[0] == scene)
    x, v = _warm(mtlp, p, x, v, 10)
    _step_vjp_check(mtlp, orc, p, x, v, scene)


@pytest.mark.parametrize("n", [1, 65, 257])
def test_step_vjp_ragged_sizes(mtlp, orc, n):
    """The backward kernels run four lanes per slot: partial waves and
    workgroups at 1, 65 and 257 particles, against the FP64 adjoint."""
    p = mtlp.scene_params(3, n)
    x, v = mtlp.generate_scene(3, n, 5, p)
    x, v = _warm(mtlp, p, x, v, 5)
    _step_vjp_check(mtlp, orc, p, x, v, f"cube_{n}")


def test_rollout_checkpointing(mtlp, orc):
    """7 steps: checkpoints every 1, 3 and 7 steps give the same gradient bit
    for bit (the recomputed segments are the steps themselves and the
    backward does not depend on atomic order), and it matches the FP64
    oracle chained over the GPU's own tapes."""
    import torch
    C = mtlp.capi
    p = mtlp.scene_params(3, 4096)
    x, v = mtlp.generate_scene(3, 4096, 9, p)
    n, steps = len(x), 7
    rng = np.random.default_rng(5)
    xbar = np.zeros((n, 4), np.float32)
    xbar[:, :3] = rng.standard_normal((n, 3))
    dev = torch.device("cuda")
    grads = {}
    with _engine(mtlp, p, x, v) as e:
        s = torch.cuda.current_stream()
        xT = None
        for every in (1, 3, 7):
            e.upload_state(x, v)
            e.rollout(DT, steps, every, s)
            xs = e.positions_buffer().read()
            if xT is None:
                xT = xs
            np.testing.assert_array_equal(xs.view(np.uint32), xT.view(np.uint32))
            xo, vo, go = (torch.empty((n, 4), device=dev), torch.empty((n, 4), device=dev),
                          torch.empty(3, device=dev))
            e.adjoint_rollout(DT, steps, every, torch.from_numpy(xbar).to(dev), None, xo, vo, go, s)
            torch.cuda.synchronize()
            grads[every] = (xo.cpu().numpy(), vo.cpu().numpy(), go.cpu().numpy())
        with pytest.raises(mtlp.MtlpError):
            e.adjoint_rollout(DT, steps, 2, None, None, None, None, None, s)  # no such rollout
        e.upload_state(x, v)
        e.tape_record(DT, steps)
        tapes = [_cpu_tape(mtlp, e, t, p.iterations) for t in range(steps)]
        assert e.error_flags() == 0
    for every in (3, 7):
        for k in range(3):
            np.testing.assert_array_equal(grads[every][k].view(np.uint32), grads[1][k].view(np.uint32),
                                          err_msg=(every, k))
    op = orc.params(p)
    rxb, rvb, rg, rg_direct = xbar.astype(np.float64), np.zeros((n, 4)), np.zeros(3), np.zeros(3)
    sums = orc.adjoint_sums(rxb, rvb)
    for t in reversed(tapes):
        dxb, dvb, g0 = orc.step_vjp_d(t, op, DT, rxb, rvb)
        rxb, rvb, g1 = orc.step_vjp_d(t, op, DT, rxb, rvb, sums)
        rg += g1
        rg_direct += g0
    errs = [_rel(grads[3][0], rxb), _rel(grads[3][1], rvb), _rel(grads[3][2], rg)]
    print("rollout 7 steps vs FP64 chain", errs, "gravity by the direct sum", _rel(grads[3][2], rg_direct),
          "growth |vbar_0|/|xbar_T|", np.linalg.norm(rvb) / np.linalg.norm(xbar))
    assert max(errs) < 1e-4, errs


def test_density_and_vjp(mtlp, orc):
    import torch
    p = mtlp.scene_params(3, 4096)
    x, v = mtlp.generate_scene(3, 4096, 2, p)
    x, v = _warm(mtlp, p, x, v, 5)
    n = len(x)
    rng = np.random.default_rng(8)
    rb = rng.standard_normal(n).astype(np.float32)
    dev = torch.device("cuda")
    with _engine(mtlp, p, x, v) as e:
        s = torch.cuda.current_stream()
        xt = torch.from_numpy(x).to(dev)
        rho = torch.empty(n, device=dev)
        xb = torch.empty((n, 4), device=dev)
        e.density(xt, rho, s)
        e.density_vjp(xt, torch.from_numpy(rb).to(dev), xb, s)
        torch.cuda.synchronize()
        rho, xb = rho.cpu().numpy(), xb.cpu().numpy()
    op = orc.params(p)
    rho_ref = orc.density_at_d(x.astype(np.float64), op)
    xb_ref = orc.density_vjp_d(x.astype(np.float64), op, rb.astype(np.float64))
    assert np.abs(rho / rho_ref - 1).max() < 1e-6
    assert _rel(xb, xb_ref) < 1e-5


def test_autograd_harness_differentiable_workload(mtlp, orc):
    """The differentiable workload (65,536 particles, h = 0.04, 3
    iterations) over 4 steps with checkpoints every 2: torch loss gradient
    w.r.t. v0 and gravity vs the FP64 oracle chained over the GPU's tapes
    (plus the oracle's density adjoint at the GPU's x_T)."""
    import torch
    from mtlphys_amd import autograd as AG
    from mtlphys_amd import workloads as W
    C = mtlp.capi
    p, x, v, target = W.differentiable()
    n, steps, every = len(x), 4, 2
    dev = torch.device("cuda")
    with _engine(mtlp, p, x, v) as e:
        x0 = torch.from_numpy(x).to(dev)
        v0 = torch.from_numpy(v).to(dev).requires_grad_(True)
        g = torch.tensor(list(p.gravity), dtype=torch.float32, requires_grad=True)
        tgt = torch.tensor(target, dtype=torch.float32, device=dev)
        xT, vT = AG.rollout(e, x0, v0, g, DT, steps, every)
        loss = AG.centroid_density_loss(e, xT, tgt)
        loss.backward()
        gv, gg = v0.grad.cpu().numpy().astype(np.float64), g.grad.numpy().astype(np.float64)
        xT_np = xT.detach().cpu().numpy()
        e.upload_state(x, v)
        e.tape_record(DT, steps)
        tapes = [_cpu_tape(mtlp, e, t, p.iterations) for t in range(steps)]
        np.testing.assert_array_equal(e.tape_read(C.TAPE_POSITIONS, steps).view(np.uint32), xT_np.view(np.uint32))
        assert e.error_flags() == 0
    op = orc.params(p)
    xTd = xT_np.astype(np.float64)
    c = xTd[:, :3].mean(axis=0) - target
    rho = orc.density_at_d(xTd, op)
    devn = rho / np.float64(op.rest_density) - 1.0
    xb = np.zeros((n, 4))
    xb[:, :3] = 2.0 * c / n
    xb += orc.density_vjp_d(xTd, op, 0.1 * 2.0 * devn / (n * np.float64(op.rest_density)))
    vb, rg = np.zeros((n, 4)), np.zeros(3)
    sums = orc.adjoint_sums(xb, vb)
    for t in reversed(tapes):
        xb, vb, g1 = orc.step_vjp_d(t, op, DT, xb, vb, sums)
        rg += g1
    L_ref = float(c @ c + 0.1 * np.mean(devn ** 2))
    print("loss", float(loss.detach()), L_ref, "dL/dv0 rel", _rel(gv, vb), "dL/dg", gg, rg)
    assert abs(float(loss.detach()) - L_ref) < 1e-4 * L_ref
    assert _rel(gv, vb) < 1e-3
    assert _rel(gg, rg) < 1e-3
    rng = np.random.default_rng(32)
    worst = 0.0
    for _ in range(32):
        d = rng.standard_normal((n, 4))
        d[:, 3] = 0
        d /= np.linalg.norm(d)
        worst = max(worst, abs(((gv - vb) * d).sum()) / np.linalg.norm(vb))
    assert worst < 1e-3, worst


def test_adjoint_rejects_batches_and_bad_calls(mtlp):
    counts = [100, 200]
    params = [mtlp.scene_params(3, c) for c in counts]
    b = mtlp.Batch(counts, params)
    b.seed(3, 1)
    with pytest.raises(mtlp.MtlpError):
        b.rollout(DT, 2, 1)
    b.close()
    p = mtlp.scene_params(3, 512)
    x, v = mtlp.generate_scene(3, 512, 1, p)
    with _engine(mtlp, p, x, v) as e:
        with pytest.raises(mtlp.MtlpError):
            e.rollout(DT, 0, 1)
        with pytest.raises(mtlp.MtlpError):
            e.rollout(0.0, 2, 1)
        with pytest.raises(mtlp.MtlpError):
            e.adjoint_rollout(DT, 2, 1)  # no rollout yet
        e.rollout(DT, 2, 1)
        with pytest.raises(mtlp.MtlpError):
            e.adjoint_rollout(DT * 2, 2, 1)
        with pytest.raises(mtlp.MtlpError):
            e.tape_read(mtlp.capi.TAPE_ITERATE, 0, 99)


def _loss_seed(orc, op, xT, target):
    """FP64 loss |centroid(x_T) - target|^2 + 0.1 mean(rho/rho0 - 1)^2 and
    its gradient w.r.t. x_T (the density term through the oracle's density
    adjoint)."""
    xTd = np.asarray(xT, np.float64)
    n = len(xTd)
    c = xTd[:, :3].mean(axis=0) - target
    rho = orc.density_at_d(xTd, op)
    devn = rho / np.float64(op.rest_density) - 1.0
    xb = np.zeros((n, 4))
    xb[:, :3] = 2.0 * c / n
    xb += orc.density_vjp_d(xTd, op, 0.1 * 2.0 * devn / (n * np.float64(op.rest_density)))
    return float(c @ c + 0.1 * np.mean(devn ** 2)), xb


def _rollout_loss_kinks(orc, op, x, v, steps, target):
    """FP64 rollout from (x, v): loss and the branch signatures of every step."""
    x, v = x.astype(np.float64), v.astype(np.float64)
    sig = []
    for _ in range(steps):
        x, v, t = orc.step_d(x, v, op, DT)
        sig.append(orc.step_kinks_d(t, op, DT))
    return _loss_seed(orc, op, x, target)[0], np.stack(sig)


def _gpu_grad(mtlp, e, p, x, v, target, steps, every):
    import torch
    from mtlphys_amd import autograd as AG
    dev = torch.device("cuda")
    e.upload_state(x, v)
    x0 = torch.from_numpy(x).to(dev)
    v0 = torch.from_numpy(v).to(dev).requires_grad_(True)
    g = torch.tensor(list(p.gravity), dtype=torch.float32, requires_grad=True)
    tgt = torch.tensor(target, dtype=torch.float32, device=dev)
    xT, _ = AG.rollout(e, x0, v0, g, DT, steps, every)
    loss = AG.centroid_density_loss(e, xT, tgt)
    loss.backward()
    return (float(loss.detach()), v0.grad.cpu().numpy().astype(np.float64), g.grad.numpy().astype(np.float64),
            xT.detach().cpu().numpy())


def test_differentiable_workload_full_length(mtlp, orc):
    """The differentiable workload as specified: 65,536 particles, h = 0.04,
    3 iterations, 40 steps with checkpoints every 8, loss |centroid(x_T) -
    target|^2 + 0.1 mean(rho/rho0 - 1)^2.  dL/dv0 and dL/dg from the torch
    harness (mtlp_rollout + mtlp_adjoint_rollout) against the FP64 oracle's
    adjoint chained over the GPU's own 40 tapes: within 1e-3 relative (norm),
    and along 32 seeded unit directions within 1e-3 of |dL/dv0|.

    Central finite differences cannot check this horizon: on the FP64 oracle
    |dL/dv0| is ~3e25 after 40 steps (~6e3 after 24), the rollout is chaotic
    and its linear range lies far below any usable step.  FD is checked where
    it is linear: the first checkpoint segment (8 steps), step 1e-4 along the
    same 32 directions: the GPU gradient must match the FP64 oracle's FD
    within 1e-3 of |dL/dv0|.  Kink crossings (particles whose branch
    signature -- bin, stencil cell, clamp, cap, wall flag, gate -- differs
    between the +-step runs in some step) are counted on the first four
    directions and reported: every direction crosses a few, with FD agreeing
    to ~1e-8 all the same, so no direction is left out."""
    from mtlphys_amd import workloads as W
    C = mtlp.capi
    p, x, v, target = W.differentiable()
    n, steps, every = len(x), W.ADJ_STEPS, W.ADJ_CKPT
    op = orc.params(p)
    print(f"differentiable workload: n={n} steps={steps} ckpt={every}", flush=True)
    with _engine(mtlp, p, x, v) as e:
        loss, gv, gg, xT = _gpu_grad(mtlp, e, p, x, v, target, steps, every)
        assert e.error_flags() == 0
        # the FP64 adjoint on the GPU's own trajectory, tape by tape from the end
        e.upload_state(x, v)
        e.tape_record(DT, steps)
        np.testing.assert_array_equal(e.tape_read(C.TAPE_POSITIONS, steps).view(np.uint32), xT.view(np.uint32))
        L_ref, xb = _loss_seed(orc, op, xT, target)
        vb, rg = np.zeros((n, 4)), np.zeros(3)
        sums = orc.adjoint_sums(xb, vb)
        for t in reversed(range(steps)):
            xb, vb, g1 = orc.step_vjp_d(_cpu_tape(mtlp, e, t, p.iterations), op, DT, xb, vb, sums)
            rg += g1
        assert e.error_flags() == 0
        # the first checkpoint segment, for the finite differences
        loss8, gv8, _, _ = _gpu_grad(mtlp, e, p, x, v, target, every, every)
        assert e.error_flags() == 0
    errs = dict(loss=abs(loss - L_ref) / L_ref, v0=_rel(gv, vb), gravity=_rel(gg, rg))
    rng = np.random.default_rng(32)
    dirs = []
    for _ in range(32):
        d = rng.standard_normal((n, 4))
        d[:, 3] = 0
        dirs.append(d / np.linalg.norm(d))
    worst = max(abs(((gv - vb) * d).sum()) / np.linalg.norm(vb) for d in dirs)
    print("40 steps vs FP64 adjoint on the GPU tapes:", errs, "worst direction", worst,
          "|dL/dv0|", np.linalg.norm(vb), "dL/dg", gg, rg, flush=True)
    assert errs["loss"] < 1e-4 and errs["v0"] < 1e-3 and errs["gravity"] < 1e-3, errs
    assert worst < 1e-3, worst

    eps = 1e-4
    fd_errs = []
    for d in dirs:
        fd = (orc.rollout_loss(x, v + eps * d, op, DT, every, target)[0] -
              orc.rollout_loss(x, v - eps * d, op, DT, every, target)[0]) / (2 * eps)
        fd_errs.append(abs(fd - (gv8 * d).sum()) / np.linalg.norm(gv8))
    # kink crossings of the first directions' +-eps runs: particles whose
    # branch signature differs in some step (they do cross, and FD is
    # unaffected at this size)
    crossing = []
    for d in dirs[:4]:
        _, sp = _rollout_loss_kinks(orc, op, x, v + eps * d, every, target)
        _, sm = _rollout_loss_kinks(orc, op, x, v - eps * d, every, target)
        crossing.append(int(np.any(sp != sm, axis=(0, 2)).sum()))
    print(f"8-step FD (step {eps}) vs GPU gradient over 32 directions: worst {max(fd_errs):.2e}, "
          f"median {np.median(fd_errs):.2e}; kink-crossing particles in directions 0-3: {crossing}", flush=True)
    assert max(fd_errs) < 1e-3, max(fd_errs)


def _remainder_pairs(mtlp, e, steps, p, per_slot_max=False):
    """Pairs (s, c) of the taped steps' passes with c in N_s (c's bin in the
    stencil of s's current cell, |r| < h) but bin(s) outside the stencil of
    c's current cell: the pairs whose share for c goes to c's inbox.  From
    the float64 neighbour pairs of a k-d tree (a count, not a parity check).
    With per_slot_max, also the most such pairs one slot receives in one
    iteration."""
    from scipy.spatial import cKDTree
    C = mtlp.capi
    dims = np.array(list(C.grid_dims(p)), np.int64)
    most = 0
    lo = np.array(p.bounds_min, np.float32)
    inv = np.float32(1.0) / np.float32(p.cell_size)
    total = 0
    for t in range(steps):
        it = [e.tape_read(C.TAPE_ITERATE, t, k)[:, :3] for k in range(p.iterations + 1)]
        cells = lambda q: np.clip(np.floor((q - lo) * inv).astype(np.int64), 0, dims - 1)
        b = cells(it[0])
        for k in range(p.iterations):
            c = cells(it[k])
            pr = cKDTree(it[k].astype(np.float64)).query_pairs(float(p.h), output_type="ndarray")
            i = np.concatenate([pr[:, 0], pr[:, 1]])
            j = np.concatenate([pr[:, 1], pr[:, 0]])
            listed = (np.abs(b[j] - c[i]) <= 1).all(1)
            rem = listed & ~(np.abs(b[i] - c[j]) <= 1).all(1)
            total += int(rem.sum())
            if rem.any():
                most = max(most, int(np.bincount(j[rem]).max()))
    return (total, most) if per_slot_max else total


def test_adjoint_deterministic_with_remainder_terms(mtlp):
    """The differentiable workload over 24 steps (checkpoints every 8), where
    thousands of pair terms reach their receiving slot through its inbox in
    the order the atomics hand out: two runs on fresh engines give the same
    bits for dL/dx0, dL/dv0 and dL/dg."""
    import torch
    from mtlphys_amd import workloads as W
    p, x, v, _ = W.differentiable()
    n, steps, every = len(x), 24, 8
    rng = np.random.default_rng(12)
    xbar = np.zeros((n, 4), np.float32)
    xbar[:, :3] = rng.standard_normal((n, 3))
    vbar = np.zeros((n, 4), np.float32)
    vbar[:, :3] = rng.standard_normal((n, 3))
    dev = torch.device("cuda")
    outs = []
    for rep in range(2):
        with _engine(mtlp, p, x, v) as e:
            s = torch.cuda.current_stream()
            e.rollout(DT, steps, every, s)
            xo, vo, go = (torch.empty((n, 4), device=dev), torch.empty((n, 4), device=dev),
                          torch.empty(3, device=dev))
            e.adjoint_rollout(DT, steps, every, torch.from_numpy(xbar).to(dev), torch.from_numpy(vbar).to(dev),
                              xo, vo, go, s)
            torch.cuda.synchronize()
            assert e.error_flags() == 0
            outs.append([xo.cpu().numpy(), vo.cpu().numpy(), go.cpu().numpy()])
            if rep == 0:
                e.upload_state(x, v)
                e.tape_record(DT, steps)
                pairs = _remainder_pairs(mtlp, e, steps, p)
    print("remainder pairs over", steps, "steps:", pairs, flush=True)
    assert pairs > 1000, pairs
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_step_vjp_dense_blob_inbox_overflow(mtlp, orc):
    """A blob packed at ~100 particles per h^3 inside a cube drop: in its
    first step the solver pushes the blob's particles the full 0.3 h cap per
    iteration, and some slots receive more remainder pair terms in one pass
    than their inbox holds (two per pair: delta and density passes), so the
    overflow path (float atomics) runs beside the inboxes.  The one-step
    gradient must still match the FP64 oracle adjoint on the GPU tape
    within 1e-4 relative (norm) per output."""
    import torch
    p = mtlp.scene_params(3, 8192)
    x, v = mtlp.generate_scene(3, 8192, 5, p)
    rng = np.random.default_rng(13)
    centre = x[:, :3].mean(axis=0)
    m = 3000
    side = (m / 100.0) ** (1.0 / 3.0)
    x[:m, :3] = (centre + rng.uniform(-0.5, 0.5, (m, 3)) * side * p.h).astype(np.float32)
    n = len(x)
    xbar = np.zeros((n, 4), np.float32)
    vbar = np.zeros((n, 4), np.float32)
    xbar[:, :3] = rng.standard_normal((n, 3))
    vbar[:, :3] = rng.standard_normal((n, 3))
    dev = torch.device("cuda")
    with _engine(mtlp, p, x, v) as e:
        s = torch.cuda.current_stream()
        e.rollout(DT, 1, 1, s)
        xo, vo, go = (torch.empty((n, 4), device=dev), torch.empty((n, 4), device=dev),
                      torch.empty(3, device=dev))
        e.adjoint_rollout(DT, 1, 1, torch.from_numpy(xbar).to(dev), torch.from_numpy(vbar).to(dev), xo, vo, go, s)
        torch.cuda.synchronize()
        gx, gv, gg = xo.cpu().numpy(), vo.cpu().numpy(), go.cpu().numpy()
        e.upload_state(x, v)
        e.tape_record(DT, 1)
        tape = _cpu_tape(mtlp, e, 0, p.iterations)
        pairs, most = _remainder_pairs(mtlp, e, 1, p, per_slot_max=True)
        assert e.error_flags() == 0
    print("remainder pairs", pairs, "most in one slot and pass", most, flush=True)
    assert 2 * most > 32, most  # the inbox of that slot overflowed
    op = orc.params(p)
    rx, rv, rg = orc.step_vjp_d(tape, op, DT, xbar.astype(np.float64), vbar.astype(np.float64),
                                orc.adjoint_sums(xbar, vbar))
    errs = dict(x=_rel(gx, rx), v=_rel(gv, rv), g=_rel(gg, rg))
    print("dense blob", errs, flush=True)
    assert errs["x"] < 1e-4 and errs["v"] < 1e-4 and errs["g"] < 1e-4, errs
