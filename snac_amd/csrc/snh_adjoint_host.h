/*
 * snh_adjoint_host.h -- host side of the adjoint (include/snac_hip.h,
 * snh_adjoint_*): checkpoints, the replay tape and the reverse sweep.
 * Included at the end of snac_hip.hip.
 *
 * Schedule (two-level checkpointing): the recorded forward run stores the
 * state entering every checkpoint_interval-th step.  The reverse sweep walks
 * the segments backwards; each segment is replayed from its checkpoint into
 * tape slots (the forward kernels read slot i and write slot i+1, so the
 * tape costs no extra traffic) and then reversed step by step.
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
    std::vector<AdjCheckpoint> ckpt;
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
};

static void adj_free(AdjointCtx* A) {
    if (!A) return;
    for (AdjCheckpoint& k : A->ckpt) {
        void* p[] = {k.stress, k.ps, k.x, k.xp, k.v, k.scal};
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
    void* p[] = {A->ts, A->tp, A->tx, A->tv, A->tf, A->tm, A->tc, A->hist, A->swork,
                 A->sbar, A->psbar, A->xbar, A->vbar, A->fbar, A->mbar, A->nb, A->grad};
    for (void* q : p)
        if (q) (void)hipFree(q);
    delete A;
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
    HIPCHK(hipMemcpyAsync(dx, sx, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dxp, sxp, 3 * N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
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
    const size_t E = c->E, N = c->N;
    k.timeStep = t;
    bool ok = hipMalloc((void**)&k.stress, 60 * E * sizeof(double)) == hipSuccess &&
              (!c->ps || hipMalloc((void**)&k.ps, 10 * E * sizeof(double)) == hipSuccess) &&
              hipMalloc((void**)&k.x, 3 * N * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&k.xp, 3 * N * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&k.v, 3 * N * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&k.scal, sizeof(DevScalars)) == hipSuccess;
    A->ckpt.push_back(k);   /* freed with the context's adjoint either way */
    if (!ok) return set_err(SNH_ENOMEM, "adjoint: cannot allocate checkpoint %zu", A->ckpt.size());
    return adj_copy_state(c, k.stress, k.ps, k.x, k.xp, k.v, k.scal, ss, sp, sx, sxp, sv, ssc);
}

extern "C" int snh_adjoint_forward(snh_context* c, int32_t nsteps) {
    Range range_("snh_adjoint_forward");
    if (!c) return set_err(SNH_EINVAL, "null context");
    AdjointCtx* A = c->adj;
    if (!A) return set_err(SNH_ESTATE, "snh_adjoint_forward before snh_adjoint_begin");
    if (nsteps < 0) return set_err(SNH_EINVAL, "negative step count");
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
        if (A->steps % A->K == 0) {
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
        const uint32_t t0 = A->ckpt[A->tape_seg].timeStep;
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
    if ((int)A->ckpt.size() != nseg) return set_err(SNH_ESTATE, "checkpoint count mismatch");
    std::vector<hipEvent_t> ev;
    for (int i = 0; i < 3 * nseg; i++) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ev.push_back(e);
    }
    bool have_nb = false;
    int rc = SNH_OK;
    for (int seg = nseg - 1; seg >= 0 && rc == SNH_OK; seg--) {
        const int s0 = seg * K;
        const int L = std::min(K, A->steps - s0);
        const AdjCheckpoint& ck = A->ckpt[seg];
        (void)hipEventRecord(ev[3 * seg], c->stream);
        /* the tape still holds the last segment recorded; others are
           replayed from their checkpoint */
        if (!(A->tape_seg == seg && A->tape_len == L)) {
            rc = adj_copy_state(c, A->ts, A->tp, A->tx + 3 * N, A->tx, A->tv, A->swork, ck.stress, ck.ps, ck.x,
                                ck.xp, ck.v, ck.scal);
            if (rc == SNH_OK) rc = adj_run_tape(c, A, ck.timeStep, 0, L);
            A->tape_seg = seg;
            A->tape_len = L;
        }
        (void)hipEventRecord(ev[3 * seg + 1], c->stream);
        /* reverse */
        for (int i = L - 1; i >= 0 && rc == SNH_OK; i--) {
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
        (void)hipEventRecord(ev[3 * seg + 2], c->stream);
    }
    if (rc == SNH_OK) {
        if (hipMemcpyAsync(grad, A->grad, E * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = set_err(SNH_EHIP, "adjoint: gradient download failed");
    } else {
        (void)hipStreamSynchronize(c->stream);
    }
    if (rc == SNH_OK) {
        for (int seg = 0; seg < nseg; seg++) {
            float a = 0.0f, b = 0.0f;
            (void)hipEventElapsedTime(&a, ev[3 * seg], ev[3 * seg + 1]);
            (void)hipEventElapsedTime(&b, ev[3 * seg + 1], ev[3 * seg + 2]);
            A->replay_ms += a;
            A->reverse_ms += b;
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
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
    out->checkpoints = (int32_t)A->ckpt.size();
    return SNH_OK;
}

#endif
