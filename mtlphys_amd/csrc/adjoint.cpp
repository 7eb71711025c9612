// adjoint.cpp -- differentiable rollout on the engine: checkpointed forward
// (mtlp_rollout), segment recompute with a tape and the backward walk
// (mtlp_adjoint_rollout), density and its adjoint for losses on the final
// state, and tape inspection.  The reference has no counterpart.
#include <cmath>
#include <cstring>

#include "engine_impl.h"

namespace {

mtlp_status check_single(const mtlp_engine* e) {
    MTLP_TRY(check_engine(e));
    if (e->batch) return fail(MTLP_E_INVALID, "the rollout adjoint runs on single-scene engines");
    return MTLP_OK;
}

// The backward stencil kernels index kAdjSub lanes per particle with 32-bit
// thread numbers.
mtlp_status check_adjoint_size(const mtlp_engine* e) {
    if (e->n > kMaxAdjointParticles) return fail(MTLP_E_RANGE, "the adjoint runs on at most 2^28 particles");
    return MTLP_OK;
}

mtlp_status check_no_sphere(const mtlp_engine* e) {
    if (e->sphere.on) return fail(MTLP_E_INVALID, "the kinematic sphere's collision is not differentiated");
    return MTLP_OK;
}

mtlp_status check_dt(float dt) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) return fail(MTLP_E_INVALID, "dt must be finite and > 0");
    return MTLP_OK;
}

size_t stride_of(const mtlp_engine* e) { return e->n_alloc; }

// Tape for `steps` steps at the engine's current shape.
mtlp_status ensure_tape(mtlp_engine* e, uint32_t steps) {
    AdjointState& A = e->adj;
    const size_t st = stride_of(e);
    const uint32_t I = e->iterations();
    if (A.tape_steps >= steps && A.stride == st && A.tape_iters == I && A.tape_cells == e->cells && A.iter)
        return MTLP_OK;
    A.release_tape();
    if (A.stride != st) {
        A.release_ckpt();
        A.release_work();
    }
    A.stride = st;
    const size_t T = steps;
    HIP_TRY(dev_alloc(&A.xs, (T + 1) * st));
    HIP_TRY(dev_alloc(&A.vs, (T + 1) * st));
    HIP_TRY(dev_alloc(&A.iter, T * (I + 1) * st + kChunkSlack));  // the stencil filter reads past the end
    HIP_TRY(dev_alloc(&A.grad, T * (I ? I : 1) * st));
    HIP_TRY(dev_alloc(&A.dlt, T * (I ? I : 1) * st));
    HIP_TRY(dev_alloc(&A.lblk, T * (I ? I : 1) * kListBlocks * st));
    HIP_TRY(dev_alloc(&A.lcnt, T * (I ? I : 1) * st));
    HIP_TRY(dev_alloc(&A.sidx, T * st));
    HIP_TRY(dev_alloc(&A.cs, T * ((size_t)e->cells + 1)));
    A.tape_steps = steps;
    A.tape_iters = I;
    A.tape_cells = e->cells;
    return MTLP_OK;
}

mtlp_status ensure_work(mtlp_engine* e) {
    AdjointState& A = e->adj;
    const size_t st = stride_of(e);
    if (A.work_stride == st && A.pacc) return MTLP_OK;
    A.release_work();
    const size_t blocks = (st + kBlock - 1) / kBlock;
    HIP_TRY(dev_alloc(&A.pacc, st));
    HIP_TRY(dev_alloc(&A.prem, st));
    HIP_TRY(dev_alloc(&A.rbar, st));
    HIP_TRY(dev_alloc(&A.db, st));
    HIP_TRY(dev_alloc(&A.lbar, st));
    HIP_TRY(dev_alloc(&A.lrem, st));
    HIP_TRY(dev_alloc(&A.rcnt, st));
    HIP_TRY(dev_alloc(&A.rval, kInbox * st));
    HIP_TRY(dev_alloc(&A.gdev, 6 * st));
    HIP_TRY(dev_alloc(&A.gpart, 6 * blocks));
    HIP_TRY(dev_alloc(&A.gstate, 9));
    HIP_TRY(dev_alloc(&A.gout, 3));
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(dev_alloc(&A.xb[k], st));
        HIP_TRY(dev_alloc(&A.vb[k], st));
        HIP_TRY(hipMemset(A.xb[k], 0, sizeof(float4) * st));
        HIP_TRY(hipMemset(A.vb[k], 0, sizeof(float4) * st));
    }
    // the remainder buffers start at zero; every backward step leaves them so
    HIP_TRY(hipMemset(A.prem, 0, sizeof(float4) * st));
    HIP_TRY(hipMemset(A.lrem, 0, sizeof(float) * st));
    HIP_TRY(hipMemset(A.rcnt, 0, sizeof(uint32_t) * st));
    A.work_stride = st;
    return MTLP_OK;
}

StepTape tape_of(const mtlp_engine* e, uint32_t slot) {
    const AdjointState& A = e->adj;
    const size_t st = A.stride;
    const uint32_t I = A.tape_iters;
    StepTape t{};
    t.sorted_index = A.sidx + slot * st;
    t.cell_start = A.cs + slot * ((size_t)A.tape_cells + 1);
    t.x_in = A.xs + slot * st;
    t.iter = A.iter + (size_t)slot * (I + 1) * st;
    t.grad = A.grad + (size_t)slot * (I ? I : 1) * st;
    t.dlt = A.dlt + (size_t)slot * (I ? I : 1) * st;
    t.lblk = A.lblk + (size_t)slot * (I ? I : 1) * kListBlocks * st;
    t.lcnt = A.lcnt + (size_t)slot * (I ? I : 1) * st;
    t.spill = e->nspill;
    t.heavy = e->heavy;
    t.heavy_count = e->heavy_count;
    t.lanes = e->lanes;
    t.stride = st;
    return t;
}

// One step from (xs[slot], vs[slot]) to (xs[slot+1], vs[slot+1]) with the
// tape of slot: the same arithmetic as enqueue_step (Engine.cpp:685-811).
// With engine_outputs it also writes the engine's lambdas (last iteration)
// and neighbour counts (iteration 0), as mtlp_step does.
mtlp_status enqueue_taped_step(mtlp_engine* e, hipStream_t s, float dt, uint32_t slot, bool engine_outputs = false) {
    AdjointState& A = e->adj;
    const StepArgs a = e->step_args(dt, true);
    const StepTape t = tape_of(e, slot);
    HashBufs b = e->hash_bufs();
    b.x = A.xs + slot * A.stride;
    b.v = A.vs + slot * A.stride;
    b.cell_start = const_cast<uint32_t*>(t.cell_start);
    b.sorted_index = const_cast<uint32_t*>(t.sorted_index);
    b.sorted = t.iterate(0);
    b.slot_of = nullptr;
    b.xyz = nullptr;  // the taped passes read the tape's float4 iterates
    StageClock clk(e, s);
    MTLP_TRY(enqueue_hash(e, s, a, clk, b));
    const uint32_t I = A.tape_iters;
    for (uint32_t it = 0; it < I; ++it) {
        float* lo = engine_outputs && it + 1 == I ? e->lambdas : nullptr;
        uint32_t* nb = engine_outputs && it == 0 ? e->nbr : nullptr;
        HIP_TRY(launch_density_lambda_tape(t.iterate(it), t.cell_start, t.sorted_index, a, t.grad_sums(it), lo, nb,
                                           t.list(it), e->stencil_block, s));
        HIP_TRY(launch_apply_delta_tape(t.iterate(it), t.cell_start, a, t.iterate(it + 1), t.delta(it), t.list(it),
                                        e->stencil_block, s));
    }
    if (engine_outputs && I == 0)
        HIP_TRY(launch_count_neighbors(t.iterate(0), t.cell_start, t.sorted_index, a, e->nbr, s));
    HIP_TRY(launch_finalize_tape(t.iterate(I), t.sorted_index, a, t.x_in, A.xs + (slot + 1) * A.stride,
                                 A.vs + (slot + 1) * A.stride, s));
    return MTLP_OK;
}

hipError_t copy_d2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
}

// Hash of d_x4 (as given, no prediction) into the engine's hash buffers.
mtlp_status enqueue_plain_hash(mtlp_engine* e, hipStream_t s, const float* d_x4) {
    const StepArgs a = e->step_args(0.0f, false);
    HashBufs b = e->hash_bufs();
    b.x = reinterpret_cast<const float4*>(d_x4);
    b.xyz = nullptr;
    StageClock clk(e, s);
    MTLP_TRY(enqueue_hash(e, s, a, clk, b));
    e->final_buf = 0;
    return MTLP_OK;
}

}  // namespace

extern "C" {

mtlp_status mtlp_rollout(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t steps, uint32_t ckpt_every) {
    MTLP_TRY(check_single(e));
    MTLP_TRY(check_adjoint_size(e));
    MTLP_TRY(check_no_sphere(e));
    MTLP_TRY(check_dt(dt));
    if (steps == 0 || ckpt_every == 0) return fail(MTLP_E_INVALID, "steps and ckpt_every must be >= 1");
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    AdjointState& A = e->adj;
    const uint32_t n_ckpt = (steps + ckpt_every - 1) / ckpt_every;
    if (A.n_ckpt < n_ckpt || A.stride != stride_of(e)) {
        HIP_TRY(hipStreamSynchronize(s));
        if (A.stride != stride_of(e)) A.release();
        A.release_ckpt();
        A.stride = stride_of(e);
        HIP_TRY(dev_alloc(&A.ck_x, (size_t)n_ckpt * A.stride));
        HIP_TRY(dev_alloc(&A.ck_v, (size_t)n_ckpt * A.stride));
        A.n_ckpt = n_ckpt;
    }
    A.recorded = false;
    // the last segment is stepped with the tape on, so the adjoint starts
    // without recomputing it
    const uint32_t last0 = (n_ckpt - 1) * ckpt_every;
    if (e->n) {
        HIP_TRY(hipStreamSynchronize(s));
        MTLP_TRY(ensure_tape(e, ckpt_every));
    }
    A.tape_last_seg = false;
    const size_t bytes = sizeof(float4) * e->n;
    for (uint32_t t = 0; t < steps; ++t) {
        if (t % ckpt_every == 0) {
            const size_t k = t / ckpt_every;
            HIP_TRY(copy_d2d(A.ck_x + k * A.stride, e->x, bytes, s));
            HIP_TRY(copy_d2d(A.ck_v + k * A.stride, e->v, bytes, s));
        }
        if (t < last0 || e->n == 0) {
            MTLP_TRY(mtlp_step(e, stream, dt));
            continue;
        }
        if (t == last0) {
            HIP_TRY(copy_d2d(A.xs, e->x, bytes, s));
            HIP_TRY(copy_d2d(A.vs, e->v, bytes, s));
        }
        MTLP_TRY(enqueue_taped_step(e, s, dt, t - last0, t + 1 == steps));
    }
    if (e->n) {
        // leave the engine's buffers as mtlp_step would
        const uint32_t k = steps - last0;
        const StepTape t = tape_of(e, k - 1);
        HIP_TRY(copy_d2d(e->x, A.xs + k * A.stride, bytes, s));
        HIP_TRY(copy_d2d(e->v, A.vs + k * A.stride, bytes, s));
        HIP_TRY(copy_d2d(e->prev, A.xs + (k - 1) * A.stride, bytes, s));
        HIP_TRY(copy_d2d(e->cell_start, t.cell_start, sizeof(uint32_t) * ((size_t)e->cells + 1), s));
        HIP_TRY(copy_d2d(e->sorted_index, t.sorted_index, sizeof(uint32_t) * e->n, s));
        HIP_TRY(copy_d2d(e->sorted[0], t.iterate(A.tape_iters), bytes, s));
        e->final_buf = 0;
        A.tape_last_seg = true;
    }
    A.recorded = true;
    A.dt = dt;
    A.steps = steps;
    A.every = ckpt_every;
    return MTLP_OK;
}

mtlp_status mtlp_adjoint_rollout(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t steps,
                                 uint32_t ckpt_every, const float* d_xbar_T, const float* d_vbar_T,
                                 float* d_xbar_0, float* d_vbar_0, float* d_gbar) {
    MTLP_TRY(check_single(e));
    AdjointState& A = e->adj;
    if (!A.recorded || A.dt != dt || A.steps != steps || A.every != ckpt_every || A.stride != stride_of(e))
        return fail(MTLP_E_INVALID, "no matching mtlp_rollout (same dt, steps, ckpt_every) before the adjoint");
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    MTLP_TRY(ensure_tape(e, ckpt_every));
    MTLP_TRY(ensure_work(e));
    const size_t bytes = sizeof(float4) * e->n;
    int cur = 0;
    if (d_xbar_T) HIP_TRY(copy_d2d(A.xb[0], d_xbar_T, bytes, s));
    else HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(A.xb[0]), 0u, 4 * (size_t)e->n, s));
    if (d_vbar_T) HIP_TRY(copy_d2d(A.vb[0], d_vbar_T, bytes, s));
    else HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(A.vb[0]), 0u, 4 * (size_t)e->n, s));
    AdjointBufs b{A.pacc, A.prem, A.rbar, A.db, A.lbar, A.lrem, A.gdev, A.gpart, A.gstate, e->misc + 2,
                  A.rcnt, A.rval};
    HIP_TRY(launch_gravity_init(b, A.xb[0], A.vb[0], e->n, A.gout, s));
    const StepArgs a = e->step_args(dt, true);
    const uint32_t n_seg = (steps + ckpt_every - 1) / ckpt_every;
    for (uint32_t seg = n_seg; seg-- > 0;) {
        const uint32_t t0 = seg * ckpt_every;
        const uint32_t t1 = t0 + ckpt_every < steps ? t0 + ckpt_every : steps;
        if (!(seg + 1 == n_seg && A.tape_last_seg)) {
            A.tape_last_seg = false;
            HIP_TRY(copy_d2d(A.xs, A.ck_x + (size_t)seg * A.stride, bytes, s));
            HIP_TRY(copy_d2d(A.vs, A.ck_v + (size_t)seg * A.stride, bytes, s));
            for (uint32_t t = t0; t < t1; ++t) MTLP_TRY(enqueue_taped_step(e, s, dt, t - t0));
        }
        for (uint32_t t = t1; t-- > t0;) {
            HIP_TRY(launch_step_vjp(tape_of(e, t - t0), b, a, A.tape_iters, e->scene_params[0].damping, A.xb[cur],
                                    A.vb[cur], A.xb[cur ^ 1], A.vb[cur ^ 1], A.gout, e->stencil_block, s));
            cur ^= 1;
        }
    }
    if (d_xbar_0) HIP_TRY(copy_d2d(d_xbar_0, A.xb[cur], bytes, s));
    if (d_vbar_0) HIP_TRY(copy_d2d(d_vbar_0, A.vb[cur], bytes, s));
    if (d_gbar) HIP_TRY(copy_d2d(d_gbar, A.gout, 3 * sizeof(float), s));
    return MTLP_OK;
}

mtlp_status mtlp_density(mtlp_engine* e, mtlp_stream stream, const float* d_x4, float* d_rho) {
    MTLP_TRY(check_single(e));
    if (e->n == 0) return MTLP_OK;
    if (!d_x4 || !d_rho) return fail(MTLP_E_INVALID, "NULL buffer");
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    MTLP_TRY(enqueue_plain_hash(e, s, d_x4));
    HIP_TRY(launch_density_rho(e->sorted[0], e->cell_start, e->sorted_index, e->step_args(0.0f, false), d_rho, s));
    return MTLP_OK;
}

mtlp_status mtlp_density_vjp(mtlp_engine* e, mtlp_stream stream, const float* d_x4, const float* d_rhobar,
                             float* d_xbar) {
    MTLP_TRY(check_single(e));
    if (e->n == 0) return MTLP_OK;
    if (!d_x4 || !d_rhobar || !d_xbar) return fail(MTLP_E_INVALID, "NULL buffer");
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    MTLP_TRY(enqueue_plain_hash(e, s, d_x4));
    HIP_TRY(launch_density_rho_vjp(e->sorted[0], e->cell_start, e->sorted_index, e->step_args(0.0f, false),
                                   d_rhobar, reinterpret_cast<float4*>(d_xbar), s));
    return MTLP_OK;
}

mtlp_status mtlp_tape_record(mtlp_engine* e, mtlp_stream stream, float dt, uint32_t steps) {
    MTLP_TRY(check_single(e));
    MTLP_TRY(check_no_sphere(e));
    MTLP_TRY(check_dt(dt));
    if (steps == 0) return fail(MTLP_E_INVALID, "steps must be >= 1");
    DeviceGuard g(e->device);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    HIP_TRY(hipStreamSynchronize(s));
    MTLP_TRY(ensure_tape(e, steps));
    AdjointState& A = e->adj;
    A.tape_last_seg = false;
    const size_t bytes = sizeof(float4) * e->n;
    HIP_TRY(copy_d2d(A.xs, e->x, bytes, s));
    HIP_TRY(copy_d2d(A.vs, e->v, bytes, s));
    for (uint32_t t = 0; t < steps; ++t) MTLP_TRY(enqueue_taped_step(e, s, dt, t));
    return MTLP_OK;
}

mtlp_status mtlp_tape_buffer(mtlp_engine* e, mtlp_tape_kind kind, uint32_t step, uint32_t iteration,
                             void** d_ptr, size_t* bytes) {
    MTLP_TRY(check_single(e));
    if (!d_ptr) return fail(MTLP_E_INVALID, "d_ptr is NULL");
    const AdjointState& A = e->adj;
    if (!A.iter) return fail(MTLP_E_INVALID, "no tape: call mtlp_tape_record first");
    const bool state = kind == MTLP_TAPE_POSITIONS || kind == MTLP_TAPE_VELOCITIES;
    if (step > A.tape_steps || (!state && step == A.tape_steps)) return fail(MTLP_E_RANGE, "step outside the tape");
    const uint32_t I = A.tape_iters;
    const StepTape t = tape_of(e, step);
    const size_t pb = sizeof(float4) * e->n;
    void* p = nullptr;
    size_t b = 0;
    switch (kind) {
        case MTLP_TAPE_POSITIONS: p = A.xs + step * A.stride; b = pb; break;
        case MTLP_TAPE_VELOCITIES: p = A.vs + step * A.stride; b = pb; break;
        case MTLP_TAPE_SORTED_INDEX: p = const_cast<uint32_t*>(t.sorted_index); b = sizeof(uint32_t) * e->n; break;
        case MTLP_TAPE_CELL_START:
            p = const_cast<uint32_t*>(t.cell_start);
            b = sizeof(uint32_t) * ((size_t)A.tape_cells + 1);
            break;
        case MTLP_TAPE_ITERATE:
            if (iteration > I) return fail(MTLP_E_RANGE, "iteration outside the tape");
            p = t.iterate(iteration);
            b = pb;
            break;
        case MTLP_TAPE_GRAD_SUMS:
        case MTLP_TAPE_DELTA:
            if (iteration >= I) return fail(MTLP_E_RANGE, "iteration outside the tape");
            p = kind == MTLP_TAPE_GRAD_SUMS ? t.grad_sums(iteration) : t.delta(iteration);
            b = pb;
            break;
        default: return fail(MTLP_E_INVALID, "unknown tape kind");
    }
    *d_ptr = p;
    if (bytes) *bytes = b;
    return MTLP_OK;
}

mtlp_status mtlp_tape_read(mtlp_engine* e, mtlp_tape_kind kind, uint32_t step, uint32_t iteration,
                           void* host_dst, size_t bytes) {
    void* p = nullptr;
    size_t b = 0;
    MTLP_TRY(mtlp_tape_buffer(e, kind, step, iteration, &p, &b));
    if (bytes > b) return fail(MTLP_E_RANGE, "read larger than the tape array");
    if (bytes == 0) return MTLP_OK;
    if (!host_dst) return fail(MTLP_E_INVALID, "host_dst is NULL");
    DeviceGuard g(e->device);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host_dst, p, bytes, hipMemcpyDeviceToHost));
    return MTLP_OK;
}

}  // extern "C"
