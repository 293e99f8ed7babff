/*
 * snh_adjoint_host.h -- host side of the adjoint (include/snac_hip.h,
 * snh_adjoint_*): checkpoints, the replay tape and the reverse sweep.
 * Included at the end of snac_hip.hip.
 *
 * The steps fall into blocks of checkpoint_interval steps.  A block is
 * reversed from a tape: its steps replayed from the state entering it into
 * tape slots (the forward kernels read slot i and write slot i+1, so the
 * tape costs no extra traffic), then reversed step by step.  Two schedules
 * provide the state entering each block:
 *
 *   uniform (snapshots == 0): the recorded forward run stores it for every
 *   block; the reverse sweep replays each block once.
 *
 *   binomial (snapshots > 0): the forward run stores only the state it
 *   starts from, and the reverse sweep follows the binomial checkpointing
 *   schedule (Griewank's revolve) over the blocks with at most `snapshots`
 *   further states held at a time: to reverse blocks [a, b) from the state
 *   entering a, it advances to a + m, keeps that state, reverses [a + m, b)
 *   with one snapshot less, then [a, a + m) with the freed one.  m minimises
 *   the blocks advanced, m + F(b - a - m, s - 1) + F(m, s), where F(n, s) =
 *   t n - C(s + 1 + t, t - 1) is the minimum for n blocks and s snapshots
 *   (t the least integer with C(s + 1 + t, t) >= n; with none, n(n - 1)/2).
 *   Memory: the tape of one block and s + 1 states, however many steps.
 *
 * Either way every step is replayed by the same kernels from the same
 * state, so both schedules give the same gradient bits.
 */
#ifndef SNH_ADJOINT_HOST_H
#define SNH_ADJOINT_HOST_H

struct AdjCheckpoint {
    double* stress = nullptr;   /* sigma* entering the step (pre-rotation) */
    double* ps = nullptr;
    double* x = nullptr;        /* coordinates of the step                  */
    double* xp = nullptr;       /* coordinates of the step before           */
    double* v = nullptr;
    DevScalars* scal = nullptr;
    uint32_t timeStep = 0;      /* timeStep of the step entered            */
};

struct AdjointCtx {
    snh_adjoint_config cfg;
    int32_t steps = 0;
    std::vector<AdjCheckpoint> ckpt;   /* uniform: one per block; binomial: the first */
    std::vector<AdjCheckpoint> snap;   /* binomial: the snapshot pool               */
    int64_t replay_steps = 0;          /* steps replayed by reverse sweeps          */
    /* tape of one segment */
    int K = 0;
    double* ts = nullptr;       /* (K+1) x [6][10][E] */
    double* tp = nullptr;       /* (K+1) x [10][E]    */
    double* tx = nullptr;       /* (K+2) x [3][N]     */
    double* tv = nullptr;       /* (K+1) x [3][N]     */
    double* tf = nullptr;       /* K x [3][N]         */
    double* tm = nullptr;       /* K x [N]            */
    uint8_t* tc = nullptr;      /* K x [N]            */
    DevScalars* hist = nullptr; /* K                  */
    DevScalars* swork = nullptr;
    /* adjoint state */
    double *sbar = nullptr, *psbar = nullptr, *xbar = nullptr, *vbar = nullptr;
    double *fbar = nullptr, *mbar = nullptr, *nb = nullptr, *grad = nullptr;
    double forward_ms = 0.0, replay_ms = 0.0, reverse_ms = 0.0;
    int tape_seg = -1;          /* segment whose steps the tape holds      */
    int tape_len = 0;           /* steps of it recorded                    */
    bool reversed = false;      /* a reverse sweep has used the tape       */
};

static void adj_free_state(AdjCheckpoint& k) {
    void* p[] = {k.stress, k.ps, k.x, k.xp, k.v, k.scal};
    for (void* q : p)
        if (q) (void)hipFree(q);
    k = AdjCheckpoint();
}

static void adj_free(AdjointCtx* A) {
    if (!A) return;
    for (AdjCheckpoint& k : A->ckpt) adj_free_state(k);
    for (AdjCheckpoint& k : A->snap) adj_free_state(k);
    void* p[] = {A->ts, A->tp, A->tx, A->tv, A->tf, A->tm, A->tc, A->hist, A->swork,
                 A->sbar, A->psbar, A->xbar, A->vbar, A->fbar, A->mbar, A->nb, A->grad};
    for (void* q : p)
        if (q) (void)hipFree(q);
    delete A;
}

/* device buffers of one state (stress, plastic strain, x, x_prev, v,
   scalars); false when one cannot be allocated (what was is kept in k) */
static bool adj_alloc_state(snh_context* c, AdjCheckpoint& k) {
    const size_t E = c->E, N = c->N;
    return hipMalloc((void**)&k.stress, 60 * E * sizeof(double)) == hipSuccess &&
           (!c->ps || hipMalloc((void**)&k.ps, 10 * E * sizeof(double)) == hipSuccess) &&
           hipMalloc((void**)&k.x, 3 * N * sizeof(double)) == hipSuccess &&
           hipMalloc((void**)&k.xp, 3 * N * sizeof(double)) == hipSuccess &&
           hipMalloc((void**)&k.v, 3 * N * sizeof(double)) == hipSuccess &&
           hipMalloc((void**)&k.scal, sizeof(DevScalars)) == hipSuccess;
}

__global__ void k_copy_scalars(const DevScalars* __restrict__ src, DevScalars* __restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

extern "C" int snh_adjoint_end(snh_context* c) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (c->adj) {
        (void)hipStreamSynchronize(c->stream);
        adj_free(c->adj);
        c->adj = nullptr;
    }
    return SNH_OK;
}

#define ADJ_ALLOC(ptr, bytes)                                                                 \
    do {                                                                                      \
        if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                               \
            adj_free(A);                                                                      \
            return set_err(SNH_ENOMEM, "adjoint: cannot allocate %zu bytes", (size_t)(bytes)); \
        }                                                                                     \
    } while (0)

extern "C" int snh_adjoint_begin(snh_context* c, const snh_adjoint_config* cfg) {
    if (!c || !cfg) return set_err(SNH_EINVAL, "null argument");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_adjoint_begin before snh_timestep_zero");
    if (c->adj) return set_err(SNH_ESTATE, "adjoint already begun (snh_adjoint_end first)");
    if (c->cfg.topo_kappa > 0)
        return set_err(SNH_EUNSUPPORTED, "the adjoint does not differentiate topography diffusion (topo_kappa > 0)");
    if (c->beta_dens)
        return set_err(SNH_EUNSUPPORTED, "the adjoint takes the tet density as constant (a material has beta != 0)");
    if (c->cfg.dtType == SNH_DT_WAVE)
        return set_err(SNH_EUNSUPPORTED, "the adjoint's mass reverse is the dynamic dt type's (dtType wave)");
    if (cfg->checkpoint_interval < 1) return set_err(SNH_EINVAL, "checkpoint_interval must be >= 1");
    if (cfg->snapshots < 0) return set_err(SNH_EINVAL, "snapshots must be >= 0");
    if (cfg->param != SNH_PARAM_LOG_VISCOSITY && cfg->param != SNH_PARAM_COHESION)
        return set_err(SNH_EINVAL, "unknown adjoint parameter %d", cfg->param);
    if (c->cfg.precision != SNH_PRECISION_FP64) return set_err(SNH_EUNSUPPORTED, "adjoint: FP64 path only");
    if (c->thermal) return set_err(SNH_EUNSUPPORTED, "adjoint: the energy entry points are not differentiated");
    if (c->cfg.dtType != SNH_DT_DYNAMIC) return set_err(SNH_EUNSUPPORTED, "adjoint: dynamic dt only");
    for (const snh_material& m : c->mats) {
        if (m.rheology & SNH_RHEOLOGY_VISCOPLASTIC)
            return set_err(SNH_EUNSUPPORTED, "adjoint: visco-plastic rheology not supported");
        if ((m.rheology & SNH_RHEOLOGY_MAXWELL) && !(m.rheology & SNH_RHEOLOGY_PLASTIC) &&
            !(m.srexponent == 1.0 && m.activationE == 0.0))
            return set_err(SNH_EUNSUPPORTED, "adjoint: Maxwell needs a linear viscosity law");
        if (m.rheology & SNH_RHEOLOGY_PLASTIC)
            for (uint32_t i = 1; i <= m.nsegments; i++)
                if (m.frictionAngle[i] != m.frictionAngle[0] || m.dilationAngle[i] != m.dilationAngle[0])
                    return set_err(SNH_EUNSUPPORTED, "adjoint: softening of friction or dilation angle");
    }
    if (cfg->param == SNH_PARAM_LOG_VISCOSITY && !c->refvisc)
        return set_err(SNH_EINVAL, "SNH_PARAM_LOG_VISCOSITY needs the SNH_FIELD_REFVISC element field");
    const size_t E = c->E, N = c->N;
    if (cfg->param == SNH_PARAM_COHESION && !c->cohoff) {
        /* a zero offset leaves every forward result unchanged */
        HIPCHK(hipMalloc((void**)&c->cohoff, E * sizeof(double)));
        HIPCHK(ctx_fill(c, c->cohoff, 0, E * sizeof(double)));
    }
    AdjointCtx* A = new AdjointCtx();
    A->cfg = *cfg;
    const int K = cfg->checkpoint_interval;
    A->K = K;
    ADJ_ALLOC(A->ts, (size_t)(K + 1) * 60 * E * sizeof(double));
    if (c->ps) ADJ_ALLOC(A->tp, (size_t)(K + 1) * 10 * E * sizeof(double));
    ADJ_ALLOC(A->tx, (size_t)(K + 2) * 3 * N * sizeof(double));
    ADJ_ALLOC(A->tv, (size_t)(K + 1) * 3 * N * sizeof(double));
    ADJ_ALLOC(A->tf, (size_t)K * 3 * N * sizeof(double));
    ADJ_ALLOC(A->tm, (size_t)K * N * sizeof(double));
    ADJ_ALLOC(A->tc, (size_t)K * N);
    ADJ_ALLOC(A->hist, (size_t)K * sizeof(DevScalars));
    ADJ_ALLOC(A->swork, sizeof(DevScalars));
    ADJ_ALLOC(A->sbar, 60 * E * sizeof(double));
    if (c->ps) ADJ_ALLOC(A->psbar, 10 * E * sizeof(double));
    ADJ_ALLOC(A->xbar, 3 * N * sizeof(double));
    ADJ_ALLOC(A->vbar, 3 * N * sizeof(double));
    ADJ_ALLOC(A->fbar, 3 * N * sizeof(double));
    ADJ_ALLOC(A->mbar, N * sizeof(double));
    ADJ_ALLOC(A->nb, E * 8 * NBC * sizeof(double));
    ADJ_ALLOC(A->grad, E * sizeof(double));
    for (int k = 0; k < cfg->snapshots; k++) {
        A->snap.emplace_back();
        if (!adj_alloc_state(c, A->snap.back())) {
            adj_free(A);
            return set_err(SNH_ENOMEM, "adjoint: cannot allocate snapshot %d of %d", k + 1, cfg->snapshots);
        }
    }
    c->adj = A;
    return SNH_OK;
}

/* run tape steps i0 .. i1-1 of the segment entered at timeStep t0: the
   forward kernels read slot i and write slot i+1; k_node also writes the
   undamped force, mass and damping branch; hist[i] keeps the scalars */
static int adj_run_tape(snh_context* c, AdjointCtx* A, uint32_t t0, int i0, int i1) {
    const size_t E = c->E, N = c->N;
    int rc = SNH_OK;
    for (int i = i0; i < i1 && rc == SNH_OK; i++) {
        const int si = sched_index(c, t0 + (uint32_t)i);
        if (si >= 0)
            hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, c->stream, A->swork, c->sched_dt + si,
                               c->sched_c + si);
        hipLaunchKernelGGL(k_copy_scalars, dim3(1), dim3(1), 0, c->stream, A->swork, A->hist + i);
        StepArgs a = make_args(c);
        a.scal = A->swork;
        a.visc = nullptr;
        a.xp = A->tx + (size_t)i * 3 * N;
        a.xc = A->tx + (size_t)(i + 1) * 3 * N;
        a.xn = A->tx + (size_t)(i + 2) * 3 * N;
        a.v = A->tv + (size_t)i * 3 * N;
        a.v_out = A->tv + (size_t)(i + 1) * 3 * N;
        a.stress = A->ts + (size_t)i * 60 * E;
        a.stress_out = A->ts + (size_t)(i + 1) * 60 * E;
        a.ps = A->tp ? A->tp + (size_t)i * 10 * E : nullptr;
        a.ps_out = A->tp ? A->tp + (size_t)(i + 1) * 10 * E : nullptr;
        if ((rc = launch_elem_args(c, a))) break;
        StepArgs an = a;
        an.stress = a.stress_out;
        an.tape_f = A->tf + (size_t)i * 3 * N;
        an.tape_m = A->tm + (size_t)i * N;
        an.tape_code = A->tc + (size_t)i * N;
        if (si >= 0 && c->sched_damp) an.damp_sched = c->sched_damp + (size_t)si * N;
        launch_node_fp64(c, an);
        StepArgs ag = a;
        ag.xc = a.xn;
        hipLaunchKernelGGL(k_geo, dim3(c->n_geo_blocks), dim3(GEO_BLOCK), 0, c->stream, ag);
        hipLaunchKernelGGL(k_fin, dim3(1), dim3(FIN_BLOCK), 0, c->stream, ag, c->n_geo_blocks, c->n_vmax_parts);
        if (hipGetLastError() != hipSuccess) rc = set_err(SNH_EHIP, "adjoint tape step launch failed");
    }
    return rc;
}

/* copy a state (stress, plastic strain, x, x_prev, v, scalars) between the
   context, a checkpoint or a tape slot */
static int adj_copy_state(snh_context* c, double* ds, double* dp, double* dx, double* dxp, double* dv,
                          DevScalars* dsc, const double* ss, const double* sp, const double* sx, const double* sxp,
                          const double* sv, const DevScalars* ssc) {
    const size_t E = c->E, N = c->N;
    HIPCHK(hipMemcpyAsync(ds, ss, 60 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (dp && sp) HIPCHK(hipMemcpyAsync(dp, sp, 10 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    /* x_prev first: rolling tape slot 1 to slot 0 reads x_prev where x goes */
    HIPCHK(hipMemcpyAsync(dxp, sxp, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dx, sx, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dv, sv, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_copy_scalars, dim3(1), dim3(1), 0, c->stream, ssc, dsc);
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

/* checkpoint of the state entering timeStep t (sources: the context's
   buffers or a tape slot) */
static int adj_save_checkpoint(snh_context* c, AdjointCtx* A, uint32_t t, const double* ss, const double* sp,
                               const double* sx, const double* sxp, const double* sv, const DevScalars* ssc) {
    AdjCheckpoint k;
    k.timeStep = t;
    const bool ok = adj_alloc_state(c, k);
    A->ckpt.push_back(k);   /* freed with the context's adjoint either way */
    if (!ok) return set_err(SNH_ENOMEM, "adjoint: cannot allocate checkpoint %zu", A->ckpt.size());
    return adj_copy_state(c, k.stress, k.ps, k.x, k.xp, k.v, k.scal, ss, sp, sx, sxp, sv, ssc);
}

/* state -> tape slot 0 */
static int adj_load(snh_context* c, AdjointCtx* A, const AdjCheckpoint& k) {
    const size_t N = c->N;
    return adj_copy_state(c, A->ts, A->tp, A->tx + 3 * N, A->tx, A->tv, A->swork, k.stress, k.ps, k.x, k.xp, k.v,
                          k.scal);
}

/* tape slot 0 -> state */
static int adj_store(snh_context* c, AdjointCtx* A, AdjCheckpoint& k, uint32_t t) {
    const size_t N = c->N;
    k.timeStep = t;
    return adj_copy_state(c, k.stress, k.ps, k.x, k.xp, k.v, k.scal, A->ts, A->tp, A->tx + 3 * N, A->tx, A->tv,
                          A->swork);
}

/* tape slot L (the state after L taped steps) -> tape slot 0 */
static int adj_roll(snh_context* c, AdjointCtx* A, int L) {
    const size_t E = c->E, N = c->N;
    if (L == 0) return SNH_OK;
    return adj_copy_state(c, A->ts, A->tp, A->tx + 3 * N, A->tx, A->tv, A->swork, A->ts + (size_t)L * 60 * E,
                          A->tp ? A->tp + (size_t)L * 10 * E : nullptr, A->tx + (size_t)(L + 1) * 3 * N,
                          A->tx + (size_t)L * 3 * N, A->tv + (size_t)L * 3 * N, A->swork);
}

extern "C" int snh_adjoint_forward(snh_context* c, int32_t nsteps) {
    Range range_("snh_adjoint_forward");
    if (!c) return set_err(SNH_EINVAL, "null context");
    AdjointCtx* A = c->adj;
    if (!A) return set_err(SNH_ESTATE, "snh_adjoint_forward before snh_adjoint_begin");
    if (nsteps < 0) return set_err(SNH_EINVAL, "negative step count");
    if (A->reversed) return set_err(SNH_ESTATE, "snh_adjoint_forward after snh_adjoint_reverse");
    if (nsteps == 0) return SNH_OK;
    if (A->steps > 0 && c->timeStep != A->ckpt[0].timeStep + (uint32_t)A->steps)
        return set_err(SNH_ESTATE, "steps were run outside snh_adjoint_forward since snh_adjoint_begin");
    const size_t E = c->E, N = c->N;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    (void)hipEventRecord(e0, c->stream);
    int rc = SNH_OK;
    int32_t done = 0;
    /* the steps run in tape mode, so the last segment needs no replay; the
       current state is in the context's buffers at entry and in tape slot
       tape_len afterwards */
    bool in_tape = false;
    while (done < nsteps && rc == SNH_OK) {
        if (A->steps % A->K == 0 && A->steps > 0 && !A->snap.empty() && in_tape) {
            /* binomial schedule: only the first state is kept */
            if ((rc = adj_roll(c, A, A->tape_len))) break;
            A->tape_seg = A->steps / A->K;
            A->tape_len = 0;
        } else if (A->steps % A->K == 0 && A->steps > 0 && !A->snap.empty()) {
            /* a new call at a block boundary: the state is in the context */
            rc = adj_copy_state(c, A->ts, A->tp, A->tx + 3 * N, A->tx, A->tv, A->swork, c->stress, c->ps,
                                c->x[c->cur], c->x[c->cur ^ 1], c->v, c->scal);
            if (rc) break;
            A->tape_seg = A->steps / A->K;
            A->tape_len = 0;
        } else if (A->steps % A->K == 0) {
            const int L = A->tape_len;
            const double* ss = in_tape ? A->ts + (size_t)L * 60 * E : c->stress;
            const double* sp = in_tape ? (A->tp ? A->tp + (size_t)L * 10 * E : nullptr) : c->ps;
            const double* sx = in_tape ? A->tx + (size_t)(L + 1) * 3 * N : c->x[c->cur];
            const double* sxp = in_tape ? A->tx + (size_t)L * 3 * N : c->x[c->cur ^ 1];
            const double* sv = in_tape ? A->tv + (size_t)L * 3 * N : c->v;
            const DevScalars* ssc = in_tape ? A->swork : c->scal;
            if ((rc = adj_save_checkpoint(c, A, c->timeStep, ss, sp, sx, sxp, sv, ssc))) break;
            const AdjCheckpoint& ck = A->ckpt.back();
            rc = adj_copy_state(c, A->ts, A->tp, A->tx + 3 * N, A->tx, A->tv, A->swork, ck.stress, ck.ps, ck.x,
                                ck.xp, ck.v, ck.scal);
            if (rc) break;
            A->tape_seg = A->steps / A->K;
            A->tape_len = 0;
        }
        in_tape = true;
        const int32_t todo = std::min(nsteps - done, A->K - A->steps % A->K);
        const uint32_t t0 = A->ckpt[0].timeStep + (uint32_t)(A->tape_seg * A->K);
        if ((rc = adj_run_tape(c, A, t0, A->tape_len, A->tape_len + todo))) break;
        A->tape_len += todo;
        done += todo;
        A->steps += todo;
        c->timeStep += todo;
    }
    if (rc == SNH_OK) {
        const int L = A->tape_len;
        rc = adj_copy_state(c, c->stress, c->ps, c->x[c->cur], c->x[c->cur ^ 1], c->v, c->scal,
                            A->ts + (size_t)L * 60 * E, A->tp ? A->tp + (size_t)L * 10 * E : nullptr,
                            A->tx + (size_t)(L + 1) * 3 * N, A->tx + (size_t)L * 3 * N, A->tv + (size_t)L * 3 * N,
                            A->swork);
        c->rotation_pending = true;
    }
    (void)hipEventRecord(e1, c->stream);
    if (rc == SNH_OK && hipEventSynchronize(e1) == hipSuccess) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        A->forward_ms += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

/* the element reverse of one step */
static int launch_elem_adj(snh_context* c, const AdjArgs& aa) {
    dim3 grid((c->E + EPW - 1) / EPW), block(64);
    switch (c->rset) {
    case RS_ELASTIC: hipLaunchKernelGGL(k_elem_adj<RS_ELASTIC>, grid, block, 0, c->stream, aa); break;
    case RS_PLASTIC: hipLaunchKernelGGL(k_elem_adj<RS_PLASTIC>, grid, block, 0, c->stream, aa); break;
    case RS_MAXWELL: hipLaunchKernelGGL(k_elem_adj<RS_MAXWELL>, grid, block, 0, c->stream, aa); break;
    /* two-material layouts: a plastic or Maxwell body with elastic parts */
    case RS_ELASTIC | RS_PLASTIC:
        hipLaunchKernelGGL(k_elem_adj<RS_ELASTIC | RS_PLASTIC>, grid, block, 0, c->stream, aa);
        break;
    case RS_ELASTIC | RS_MAXWELL:
        hipLaunchKernelGGL(k_elem_adj<RS_ELASTIC | RS_MAXWELL>, grid, block, 0, c->stream, aa);
        break;
    default:
        hipLaunchKernelGGL(k_elem_adj<RS_ELASTIC | RS_PLASTIC | RS_MAXWELL>, grid, block, 0, c->stream, aa);
        break;
    }
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

/* blocks advanced by the binomial schedule to reverse n blocks from a
   kept state with s snapshots free: t n - C(s + 1 + t, t - 1), t the least
   integer with C(s + 1 + t, t) >= n (Griewank's count with s + 1 states) */
static int64_t binomial_cost(int64_t n, int64_t s) {
    if (n <= 1) return 0;
    const double S = (double)(s + 1);
    int64_t t = 0;
    double b = 1.0;   /* C(S + t, t) */
    while (b < (double)n) {
        t++;
        b = b * (S + (double)t) / (double)t;
    }
    return t * n - (int64_t)llround(b * (double)t / (S + 1.0));
}

/* the split m of [a, a + n) with s snapshots free: advance m blocks, keep
   that state, reverse the n - m blocks after it with s - 1, then the m
   before it with s */
static int64_t binomial_split(int64_t n, int64_t s) {
    int64_t best = 1, cost = -1;
    for (int64_t m = 1; m < n; m++) {
        const int64_t v = m + binomial_cost(n - m, s - 1) + binomial_cost(m, s);
        if (cost < 0 || v < cost) {
            cost = v;
            best = m;
        }
    }
    return best;
}

/* one reverse sweep: tapes, advances and reverses of blocks */
struct AdjSweep {
    snh_context* c;
    AdjointCtx* A;
    std::vector<std::pair<hipEvent_t, int>> marks;   /* (event, 0 replay / 1 reverse / 2 end) */
    bool have_nb;

    int len(int blk) const { return std::min(A->K, A->steps - blk * A->K); }
    uint32_t t0(int blk) const { return A->ckpt[0].timeStep + (uint32_t)(blk * A->K); }
    void mark(int kind) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        (void)hipEventRecord(e, c->stream);
        marks.push_back({e, kind});
    }
    /* tape slot 0 holds the state entering blk: replay the block into the tape */
    int tape(int blk) {
        const int L = len(blk);
        const int rc = adj_run_tape(c, A, t0(blk), 0, L);
        A->replay_steps += L;
        A->tape_seg = blk;
        A->tape_len = L;
        return rc;
    }
    /* tape slot 0 holds the state entering block `from`: replay blocks up to
       `to`, after which slot 0 holds the state entering `to` */
    int advance(int from, int to) {
        int rc = SNH_OK;
        for (int b = from; b < to && rc == SNH_OK; b++) {
            rc = tape(b);
            if (rc == SNH_OK) rc = adj_roll(c, A, len(b));
        }
        A->tape_seg = -1;
        A->tape_len = 0;
        return rc;
    }
    /* reverse the steps of blk, whose replay the tape holds */
    int reverse(int blk) {
        const size_t E = c->E, N = c->N;
        int rc = SNH_OK;
        for (int i = len(blk) - 1; i >= 0 && rc == SNH_OK; i--) {
            NodeAdjArgs na;
            na.nb = have_nb ? A->nb : nullptr;
            na.xbar = A->xbar;
            na.vbar = A->vbar;
            na.fbar = A->fbar;
            na.mbar = A->mbar;
            na.tape_f = A->tf + (size_t)i * 3 * N;
            na.tape_m = A->tm + (size_t)i * N;
            na.tape_code = A->tc + (size_t)i * N;
            na.bc = c->bc;
            na.bcs = c->dbc;
            na.scal = A->hist + i;
            na.demf = c->cfg.demf;
            na.I = c->I; na.J = c->J; na.K = c->K;
            na.EI = c->EI; na.EJ = c->EJ; na.EK = c->EK;
            na.N = c->N; na.E = c->E;
            hipLaunchKernelGGL(k_node_adj, dim3(c->n_node_blocks), dim3(NODE_BLOCK), 0, c->stream, na);
            AdjArgs aa;
            aa.f = make_args(c);
            aa.f.scal = A->hist + i;
            aa.f.visc = nullptr;
            aa.f.xp = A->tx + (size_t)i * 3 * N;
            aa.f.xc = A->tx + (size_t)(i + 1) * 3 * N;
            aa.f.v = A->tv + (size_t)i * 3 * N;
            aa.f.stress = A->ts + (size_t)i * 60 * E;
            aa.f.stress_out = A->ts + (size_t)(i + 1) * 60 * E;   /* sigma*_n of the replay */
            aa.f.ps = A->tp ? A->tp + (size_t)i * 10 * E : nullptr;
            aa.sbar = A->sbar;
            aa.psbar = A->psbar;
            aa.fbar = A->fbar;
            aa.mbar = A->mbar;
            aa.grad = A->grad;
            aa.nb = A->nb;
            aa.param = A->cfg.param;
            aa.flags = A->cfg.flags;
            rc = launch_elem_adj(c, aa);
            have_nb = true;
        }
        return rc;
    }
    /* reverse blocks [a, b) from ck, the state entering a, with the
       snapshots whose indices pool holds (given back on return) */
    int revolve(int a, int b, const AdjCheckpoint& ck, std::vector<int>& pool) {
        int rc = SNH_OK;
        if (b - a == 1 || pool.empty()) {
            for (int j = b - 1; j >= a && rc == SNH_OK; j--) {
                mark(0);
                rc = adj_load(c, A, ck);
                if (rc == SNH_OK) rc = advance(a, j);
                if (rc == SNH_OK) rc = tape(j);
                mark(1);
                if (rc == SNH_OK) rc = reverse(j);
            }
            return rc;
        }
        const int m = (int)binomial_split(b - a, (int64_t)pool.size());
        const int slot = pool.back();
        pool.pop_back();
        mark(0);
        rc = adj_load(c, A, ck);
        if (rc == SNH_OK) rc = advance(a, a + m);
        if (rc == SNH_OK) rc = adj_store(c, A, A->snap[slot], t0(a + m));
        if (rc == SNH_OK) rc = revolve(a + m, b, A->snap[slot], pool);
        pool.push_back(slot);
        if (rc == SNH_OK) rc = revolve(a, a + m, ck, pool);
        return rc;
    }
};

extern "C" int snh_adjoint_reverse(snh_context* c, const double* coord_bar, const double* vel_bar, double* grad) {
    Range range_("snh_adjoint_reverse");
    if (!c || !grad) return set_err(SNH_EINVAL, "null argument");
    AdjointCtx* A = c->adj;
    if (!A) return set_err(SNH_ESTATE, "snh_adjoint_reverse before snh_adjoint_begin");
    if (A->steps == 0) return set_err(SNH_ESTATE, "no recorded steps");
    const size_t E = c->E, N = c->N;
    const int K = A->K;
    /* seeds: node-major [N][3] -> [3][N] */
    std::vector<double> seed(3 * N, 0.0);
    HIPCHK(hipMemsetAsync(A->sbar, 0, 60 * E * sizeof(double), c->stream));
    if (A->psbar) HIPCHK(hipMemsetAsync(A->psbar, 0, 10 * E * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(A->grad, 0, E * sizeof(double), c->stream));
    /* components 6..8 of nb are written only where a Maxwell law is compiled in */
    HIPCHK(hipMemsetAsync(A->nb, 0, E * 8 * NBC * sizeof(double), c->stream));
    for (int which = 0; which < 2; which++) {
        const double* src = which == 0 ? coord_bar : vel_bar;
        double* dst = which == 0 ? A->xbar : A->vbar;
        if (src) {
            for (size_t n = 0; n < N; n++)
                for (int d = 0; d < 3; d++) seed[d * N + n] = src[3 * n + d];
            HIPCHK(hipMemcpyAsync(dst, seed.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));   /* seed is reused */
        } else {
            HIPCHK(hipMemsetAsync(dst, 0, 3 * N * sizeof(double), c->stream));
        }
    }
    const int nseg = (A->steps + K - 1) / K;
    const bool binomial = !A->snap.empty();
    if ((int)A->ckpt.size() != (binomial ? 1 : nseg)) return set_err(SNH_ESTATE, "checkpoint count mismatch");
    AdjSweep sw{c, A, {}, false};
    A->reversed = true;
    int rc = SNH_OK;
    if (!binomial) {
        for (int seg = nseg - 1; seg >= 0 && rc == SNH_OK; seg--) {
            sw.mark(0);
            /* the tape still holds the last block recorded; others are
               replayed from their checkpoint */
            if (!(A->tape_seg == seg && A->tape_len == sw.len(seg))) {
                rc = adj_load(c, A, A->ckpt[seg]);
                if (rc == SNH_OK) rc = sw.tape(seg);
            }
            sw.mark(1);
            if (rc == SNH_OK) rc = sw.reverse(seg);
        }
    } else {
        int nrev = nseg;
        if (A->tape_seg == nseg - 1 && A->tape_len == sw.len(nseg - 1)) {
            sw.mark(1);
            rc = sw.reverse(nseg - 1);
            nrev = nseg - 1;
        }
        std::vector<int> pool;
        for (int k = (int)A->snap.size() - 1; k >= 0; k--) pool.push_back(k);
        if (rc == SNH_OK && nrev > 0) rc = sw.revolve(0, nrev, A->ckpt[0], pool);
    }
    sw.mark(2);
    if (rc == SNH_OK) {
        if (hipMemcpyAsync(grad, A->grad, E * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = set_err(SNH_EHIP, "adjoint: gradient download failed");
    } else {
        (void)hipStreamSynchronize(c->stream);
    }
    if (rc == SNH_OK) {
        /* marks: 0 opens a replay, 1 a reverse, 2 closes the last phase */
        for (size_t k = 0; k + 1 < sw.marks.size(); k++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, sw.marks[k].first, sw.marks[k + 1].first);
            (sw.marks[k].second == 0 ? A->replay_ms : A->reverse_ms) += ms;
        }
    }
    for (auto& m : sw.marks) (void)hipEventDestroy(m.first);
    return rc;
}

extern "C" int snh_adjoint_get_times(snh_context* c, snh_adjoint_times* out) {
    if (!c || !out) return set_err(SNH_EINVAL, "null argument");
    AdjointCtx* A = c->adj;
    if (!A) return set_err(SNH_ESTATE, "no adjoint begun");
    out->forward_ms = A->forward_ms;
    out->replay_ms = A->replay_ms;
    out->reverse_ms = A->reverse_ms;
    out->steps = A->steps;
    out->checkpoints = (int32_t)(A->snap.empty() ? A->ckpt.size() : A->snap.size() + 1);
    out->replay_steps = A->replay_steps;
    return SNH_OK;
}

#endif
