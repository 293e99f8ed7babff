/*
 * snh_mixed.h -- SNH_PRECISION_MIXED kernels: strain-rate, rotation,
 * constitutive-increment and force arithmetic in FP32; node coordinates,
 * velocities, tet stresses and plastic strain stay FP64.
 *
 * Per-step increments are ~1e-8 of the state they update, so every FP32
 * quantity is an increment, a geometry term built from FP64 coordinate
 * differences, or a difference of stresses taken in FP64 first:
 *   - element geometry: edge vectors relative to the tet's first node are
 *     formed in FP64 and rounded once to FP32;
 *   - stress rotation and the elastic / Maxwell updates are FP32 increments
 *     added to the FP64 stress;
 *   - trace averaging and the mean stress stay FP64;
 *   - Maxwell's volumetric strain V/V_old - 1 (a difference of nearly equal
 *     volumes) is formed in FP64;
 *   - the Mohr-Coulomb eigen-solve and return run in FP64 on the FP64 trial
 *     stress;
 *   - the nodal force is gathered relative to a reference stress per
 *     surrounding element (its first tet): the reference part
 *     sigma_ref . sum(n A)/6 is exact FP64, the remainder, small stress
 *     differences times FP32 face vectors, is summed in FP64.
 * Within one build the result is deterministic (atomic-free, fixed order).
 */
#ifndef SNH_MIXED_H
#define SNH_MIXED_H

struct TetKinF {
    float V;
    float sr[6];
    float w[3];
};

/* E: tet node coordinates relative to its node 0 (E[0] = 0), U: velocities.
   Uses n A = a / 2 with a the face cross product (the identity behind
   StrainRate.c:65-184), so no face needs a square root or a division:
   L = -(1 / 6V) sum_k u_k (x) a_(k+1)%4, strain rate sym(L), spin skew(L). */
__device__ __forceinline__ void tet_kinematics_f(const float E[4][3], const float U[4][3], TetKinF& o) {
#pragma clang fp contract(fast)
    float a[4][3];
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const float* p0 = E[kFaceToNode[f][0]];
        const float* p1 = E[kFaceToNode[f][1]];
        const float* p2 = E[kFaceToNode[f][2]];
        const float x2 = p1[0] - p0[0], y2 = p1[1] - p0[1], z2 = p1[2] - p0[2];
        const float x3 = p2[0] - p0[0], y3 = p2[1] - p0[1], z3 = p2[2] - p0[2];
        a[f][0] = y2 * z3 - z2 * y3;
        a[f][1] = z2 * x3 - x2 * z3;
        a[f][2] = x2 * y3 - y2 * x3;
    }
    const float V = fabsf((E[3][0] * a[0][0] + E[3][1] * a[0][1] + E[3][2] * a[0][2]) * (1.0f / 6.0f));
    o.V = V;
    float G[3][3];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
            G[p][q] = U[0][p] * a[1][q] + U[1][p] * a[2][q] + U[2][p] * a[3][q] + U[3][p] * a[0][q];
    const float g = -__builtin_amdgcn_rcpf(6.0f * V);   /* L = g G */
    o.sr[0] = g * G[0][0];
    o.sr[1] = g * G[1][1];
    o.sr[2] = g * G[2][2];
    o.sr[3] = 0.5f * g * (G[0][1] + G[1][0]);
    o.sr[4] = 0.5f * g * (G[0][2] + G[2][0]);
    o.sr[5] = 0.5f * g * (G[1][2] + G[2][1]);
    o.w[0] = 0.5f * g * (G[0][1] - G[1][0]);
    o.w[1] = 0.5f * g * (G[0][2] - G[2][0]);
    o.w[2] = 0.5f * g * (G[1][2] - G[2][1]);
}

/* Jaumann rotation as FP32 increments applied in the reference's order to
   the FP64 stress (UpdateElement.c:245-265) */
__device__ __forceinline__ void rotate_stress_mx(double st[6], const float w[3], float dt) {
#pragma clang fp contract(fast)
    const float w01 = w[0] * dt, w02 = w[1] * dt, w12 = w[2] * dt;
    st[0] += (double)(2.0f * (w01 * (float)st[3] + w02 * (float)st[4]));
    st[1] += (double)(2.0f * (-w01 * (float)st[3] + w12 * (float)st[5]));
    st[2] += (double)(2.0f * (-w02 * (float)st[4] - w12 * (float)st[5]));
    st[3] += (double)(w01 * (float)(st[1] - st[0]) + w02 * (float)st[5] + w12 * (float)st[4]);
    st[4] += (double)(w02 * (float)(st[2] - st[0]) + w01 * (float)st[5] - w12 * (float)st[3]);
    st[5] += (double)(w12 * (float)(st[2] - st[1]) - w01 * (float)st[4] - w02 * (float)st[3]);
}

/* TW as in k_elem: one wave per tet index over 64 elements */
template <int RSET, bool TW>
__device__ __forceinline__ void elem_pass_mx(StepArgs a, const int blk, const int nblk) {
    constexpr int NE = TW ? 64 : EPW;
    constexpr int NT = TW ? NTET + 1 : NTET;
    __shared__ double sh_a[NE][NT];
    __shared__ double sh_b[NE][NT];
    __shared__ double sh_c[NE][NT];   /* the stress-trace sums */
    __shared__ double sh_sp[NE][NT][5];
    __shared__ double sh_ps[NE][NT];
    const int lane = threadIdx.x;
    const bool in_wave = TW || lane < EPW * NTET;
    const int le = TW ? (lane & 63) : (in_wave ? lane / NTET : 0);
    const int t = TW ? (lane >> 6) : (in_wave ? lane - le * NTET : 0);
    const int e = xcd_block(blk, nblk) * NE + le;
    const bool active = in_wave && e < a.E;
    const size_t E = (size_t)a.E;
    const int ec = active ? e : 0;
    int ei, ej, ek;
    elem_ijk(a, ec, ei, ej, ek);
    const double dt = a.scal->dt;
    const float dtf = (float)dt;
    const float dt_rotf = (float)a.scal->dt_rot;
    const DevMaterial& m = a.mats[a.mat ? a.mat[ec] : 0];
    /* the element's viscosity, loaded ahead of the kinematics and the block
       barriers so that its latency overlaps them (1.5 % off the Maxwell pass) */
    const double rvisc_e = (RSET & RS_MAXWELL) && a.refvisc ? a.refvisc[ec] : m.refvisc;
    /* and the stored plastic strain */
    const bool is_pl = (RSET & RS_PLASTIC) && (m.rheology & SNH_RHEOLOGY_PLASTIC);
    double ps = 0.0;
    if ((RSET & RS_PLASTIC) && is_pl && active) ps = a.ps[(size_t)t * E + ec];

    int nd[4];
    double X0[3];
    float Ef[4][3], Uf[4][3];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        nd[q] = hex_node_index(a, ei, ej, ek, tet_node(t, q));
        const double x = a.xc[nd[q]], y = a.xc[a.N + nd[q]], z = a.xc[2 * a.N + nd[q]];
        if (q == 0) {
            X0[0] = x;
            X0[1] = y;
            X0[2] = z;
        }
        Ef[q][0] = (float)(x - X0[0]);
        Ef[q][1] = (float)(y - X0[1]);
        Ef[q][2] = (float)(z - X0[2]);
        Uf[q][0] = (float)a.v[nd[q]];
        Uf[q][1] = (float)a.v[a.N + nd[q]];
        Uf[q][2] = (float)a.v[2 * a.N + nd[q]];
    }
    TetKinF k;
    tet_kinematics_f(Ef, Uf, k);
    const float trace = (k.sr[0] + k.sr[1] + k.sr[2]) * (1.0f / 3.0f);
    if (in_wave) {
        sh_a[le][t] = (double)k.V;
        sh_b[le][t] = (double)(trace * k.V);
    }
    __syncthreads();
    double evol = 0.0, partial = 0.0;
#pragma unroll
    for (int q = 0; q < NTET; q++) {
        evol += sh_a[le][q];
        partial += sh_b[le][q];
    }
    const double elemVolume = evol;
    const double rvol = 1.0 / elemVolume;
    const float srcorr = (float)(partial * rvol);

    double st[6];
#pragma unroll
    for (int c = 0; c < 6; c++) st[c] = active ? a.stress[((size_t)c * NTET + t) * E + ec] : 0.0;
    rotate_stress_mx(st, k.w, dt_rotf);
    float sn[6];
#pragma unroll
    for (int c = 0; c < 3; c++) sn[c] = (k.sr[c] - trace + srcorr) * dtf;
#pragma unroll
    for (int c = 3; c < 6; c++) sn[c] = k.sr[c] * dtf;

    Soft sp = {0.0, 0.0, 0.0, 0.0};
    if (RSET & RS_PLASTIC) {
        double matched = 0.0;
        Soft own = {0.0, 0.0, 0.0, 0.0};
        if (is_pl) {
            for (uint32_t i = 0; i < m.nseg; i++) {
                const double pl1 = m.plstrain[i], pl2 = m.plstrain[i + 1];
                if (ps >= pl1 && ps <= pl2) {
                    const double tgf = m.tgf[i];
                    const double tgd = m.tgd[i];
                    const double tgc = m.tgc[i];
                    own.phi = m.fa[i] + tgf * (ps - pl1);
                    own.psi = m.da[i] + tgd * (ps - pl1);
                    own.coh = m.coh[i] + tgc * (ps - pl1);
                    own.hard = tgc;
                    matched = 1.0;
                }
            }
        }
        if (in_wave) {
            sh_sp[le][t][0] = matched;
            sh_sp[le][t][1] = own.phi;
            sh_sp[le][t][2] = own.psi;
            sh_sp[le][t][3] = own.coh;
            sh_sp[le][t][4] = own.hard;
            sh_ps[le][t] = ps;
        }
        __syncthreads();
        if (is_pl) {
            sp.phi = 30.0;
            sp.coh = 30.0e6;
            for (int q = 0; q <= t; q++) {
                if (sh_sp[le][q][0] != 0.0) {
                    sp.phi = sh_sp[le][q][1];
                    sp.psi = sh_sp[le][q][2];
                    sp.coh = sh_sp[le][q][3];
                    sp.hard = sh_sp[le][q][4];
                }
                if (!(sp.phi >= 0.0) && sh_ps[le][q] < 0.0) {
                    sp.phi = m.fa[0];
                    sp.psi = m.da[0];
                    sp.coh = m.coh[0];
                    if (q == t) ps = 0.0;
                }
            }
        }
    }

    if (is_pl) {
        const float trs = sn[0] + sn[1] + sn[2];
        const float twomu = (float)(2.0 * m.mu), lam = (float)m.lambda;
        st[0] += (double)(twomu * sn[0] + lam * trs);
        st[1] += (double)(twomu * sn[1] + lam * trs);
        st[2] += (double)(twomu * sn[2] + lam * trs);
        st[3] += (double)(twomu * sn[3]);
        st[4] += (double)(twomu * sn[4]);
        st[5] += (double)(twomu * sn[5]);
        double S[3][3];
        S[0][0] = st[0]; S[1][1] = st[1]; S[2][2] = st[2];
        S[0][1] = S[1][0] = st[3];
        S[0][2] = S[2][0] = st[4];
        S[1][2] = S[2][1] = st[5];
        const double coh = a.cohoff ? sp.coh + a.cohoff[ec] : sp.coh;
        double tc = m.ten_off;
        if (sp.phi > 0.0 && tc == 0.0) tc = coh / tan_phi(m, sp.phi);
        double sphi, spsi, dps = 0.0;
        angle_terms(m, sp.phi, sp.psi, sphi, spsi);
        if (!mc_surely_elastic(m, st, sphi, coh, tc)) {
            double d3[3], cn[3][3];
            principal_stresses(S, d3, cn);
            if (mc_return(m, S, d3, cn, sphi, spsi, coh, tc, 0.0, false, dps)) {
                ps = ps + dps;
                st[0] = S[0][0]; st[1] = S[1][1]; st[2] = S[2][2];
                st[3] = S[0][1]; st[4] = S[0][2]; st[5] = S[1][2];
            }
        }
        if (active) a.ps_out[(size_t)t * E + ec] = ps;
    } else if ((RSET & RS_MAXWELL) && (m.rheology & SNH_RHEOLOGY_MAXWELL)) {
        /* volumetric strain from FP64 volumes of the two configurations */
        double Xc[4][3], Xp[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Xc[q][0] = a.xc[nd[q]];
            Xc[q][1] = a.xc[a.N + nd[q]];
            Xc[q][2] = a.xc[2 * a.N + nd[q]];
            Xp[q][0] = a.xp[nd[q]];
            Xp[q][1] = a.xp[a.N + nd[q]];
            Xp[q][2] = a.xp[2 * a.N + nd[q]];
        }
        const double v_new = tet_volume(Xc[0], Xc[1], Xc[2], Xc[3]);
        const double v_old = tet_volume(Xp[0], Xp[1], Xp[2], Xp[3]);
        /* a difference of nearly equal FP64 volumes, then an FP32 quotient */
        const double trace_strain = (double)((float)(v_new - v_old) / (float)v_old);
        add_thermal_stress(a, st, t, ec);
        const double rvisc = rvisc_e;
        double visc;
        {
            const double te3 = trace_strain * (1.0 / 3.0);
            double d0 = sn[0] - te3, d1 = sn[1] - te3, d2 = sn[2] - te3;
            const double snd[6] = {sn[0], sn[1], sn[2], sn[3], sn[4], sn[5]};
            visc = viscosity_of(m, rvisc, d0, d1, d2, snd, dt, tet_temp(a, m, nd));
        }
        if (a.visc && active) a.visc[(size_t)t * E + ec] = visc;
        /* sigma'_new = sigma' + sigma'(-2x/(1+x)) + 2 mu e'/(1+x), x = mu dt/(2 eta) */
        const double mean = (st[0] + st[1] + st[2]) * (1.0 / 3.0);
        const float x = (float)m.mu / (2.0f * (float)visc) * dtf;
        const float fdec = -2.0f * x / (1.0f + x);
        const float gmu = (float)(2.0 * m.mu) / (1.0f + x);
        const float te3f = (float)trace_strain * (1.0f / 3.0f);
        const double bulk_inc = (m.lambda + 2.0 * m.mu / 3.0) * trace_strain;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float sd = (float)(st[c] - mean);
            st[c] += (double)(sd * fdec + gmu * (sn[c] - te3f)) + bulk_inc;
        }
#pragma unroll
        for (int c = 3; c < 6; c++) st[c] += (double)((float)st[c] * fdec + gmu * sn[c]);
    } else if ((RSET & RS_ELASTIC) && (m.rheology & SNH_RHEOLOGY_ELASTIC)) {
        const float trs = sn[0] + sn[1] + sn[2];
        const float twomu = (float)(2.0 * m.mu), lam = (float)m.lambda;
        st[0] += (double)(twomu * sn[0] + lam * trs);
        st[1] += (double)(twomu * sn[1] + lam * trs);
        st[2] += (double)(twomu * sn[2] + lam * trs);
        st[3] += (double)(twomu * sn[3]);
        st[4] += (double)(twomu * sn[4]);
        st[5] += (double)(twomu * sn[5]);
    }

    /* stress trace averaging in FP64 (Stress.c:87-137) */
    const double tr = (st[0] + st[1] + st[2]) * (1.0 / 3.0);
    /* its own array: no wait for the readers of sh_b */
    if (in_wave) sh_c[le][t] = tr * (double)k.V;
    __syncthreads();
    double pst = 0.0;
#pragma unroll
    for (int q = 0; q < NTET; q++) pst += sh_c[le][q];
    const double avgTrace = pst * rvol;
    st[0] += -tr + avgTrace;
    st[1] += -tr + avgTrace;
    st[2] += -tr + avgTrace;
    if (active) {
#pragma unroll
        for (int c = 0; c < 6; c++) a.stress_out[((size_t)c * NTET + t) * E + ec] = st[c];
        if (a.tpre) store_density(a, m, nd, t, ec);
    }
}

template <int RSET, bool TW>
__global__ void __launch_bounds__(TW ? 640 : 64) k_elem_mx(StepArgs a) {
    elem_pass_mx<RSET, TW>(a, (int)blockIdx.x, (int)gridDim.x);
}

/* node gather, mixed precision */
__device__ __forceinline__ void node_pass_mx(StepArgs a, const int blk, const int nblk) {
    __shared__ double sh[NODE_BLOCK / 64];
    const int g = xcd_block(blk, nblk) * blockDim.x + threadIdx.x;
    double vmag = -INFINITY;
    if (g < a.N) {
        /* the FP64 gather's banded node order: node rows that read one
           element row run close together and share its L2 */
        int i, j, k;
        const int n = gather_node(a, g, i, j, k);
        const double dt = a.scal->dt;
        const double c = a.scal->speedOfSound;
        const double rc2 = 1.0 / (c * c);
        const size_t E = (size_t)a.E;
        const NodeOwn own = node_own(a, n);
        const double xn0 = own.x0, xn1 = own.x1, xn2 = own.x2;
        double im = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
        for (int slot = 0; slot < 8; slot++) {
            const int ei = i - 1 + (slot & 1), ej = j - 1 + ((slot >> 1) & 1), ek = k - 1 + ((slot >> 2) & 1);
            if (ei < 0 || ej < 0 || ek < 0 || ei >= a.EI || ej >= a.EJ || ek >= a.EK) continue;
            const int e = ei + a.EI * (ej + a.EJ * ek);
            const DevMaterial& m = a.mats[a.mat ? a.mat[e] : 0];
            double X[8][3];
            load_hex(a, a.xc, ei, ej, ek, X);
            /* reference stress of this element: its first listed tet */
            const int tref = kNodeToTetra[slot][0];
            double sref[6];
#pragma unroll
            for (int q = 0; q < 6; q++) sref[q] = a.stress[((size_t)q * NTET + tref) * E + e];
            double C0 = 0.0, C1 = 0.0, C2 = 0.0;   /* sum of FP64 face vectors (2 n A) */
            double vs = 0.0;                         /* sum of the 5 tet volumes, FP64 */
            double vr = 0.0;                         /* with tet densities: sum of rho V */
            float g0 = 0.f, g1 = 0.f, g2 = 0.f;
#pragma unroll
            for (int idx = 0; idx < 5; idx++) {
#pragma clang fp contract(fast)
                const int t = kNodeToTetra[slot][idx];
                const int f = kNodeToSurface[slot][idx];
                const int q0 = kTetraToNode[t][kFaceToNode[f][0]];
                const int q1 = kTetraToNode[t][kFaceToNode[f][1]];
                const int q2 = kTetraToNode[t][kFaceToNode[f][2]];
                /* the tet's fourth node, opposite face f */
                const int qo = kTetraToNode[t][6 - kFaceToNode[f][0] - kFaceToNode[f][1] - kFaceToNode[f][2]];
                double dax, day, daz;
                face_cross(X[q0], X[q1], X[q2], dax, day, daz);
                C0 += dax;
                C1 += day;
                C2 += daz;
                /* tet volume from the same face: |(x_o - x_0) . (2 n A)| / 6.
                   FP64: the gravity load it scales is balanced by the
                   FP64 pre-stress term, so its rounding error is not
                   damped by the force balance. */
                const double vt = fabs((X[qo][0] - X[q0][0]) * dax + (X[qo][1] - X[q0][1]) * day +
                                       (X[qo][2] - X[q0][2]) * daz);
                vs += vt;
                if (a.dens) vr += a.dens[(size_t)t * E + e] * vt;
                const float ax = (float)dax, ay = (float)day, az = (float)daz;   /* 2 n A */
                float ds[6];
#pragma unroll
                for (int q = 0; q < 6; q++) ds[q] = (float)(a.stress[((size_t)q * NTET + t) * E + e] - sref[q]);
                g0 += ds[0] * ax + ds[3] * ay + ds[4] * az;
                g1 += ds[3] * ax + ds[1] * ay + ds[5] * az;
                g2 += ds[4] * ax + ds[5] * ay + ds[2] * az;
            }
            vs *= 1.0 / 6.0;
            /* the tet densities of the temperature plugin's last Stress call,
               or the material's */
            const double gr = a.gravity * 0.125 * (a.dens ? vr * (1.0 / 6.0) : m.rho_eff * vs);
            const double mi = 0.25 * (m.lam2mu * rc2) * vs;
            /* sigma . n A / 6 = sigma . (2 n A) / 12 */
            constexpr double r12 = 1.0 / 12.0;
            f0 += (sref[0] * C0 + sref[3] * C1 + sref[4] * C2) * r12 + (double)g0 * r12;
            f1 += (sref[3] * C0 + sref[1] * C1 + sref[5] * C2) * r12 + (double)g1 * r12 - gr;
            f2 += (sref[4] * C0 + sref[5] * C1 + sref[2] * C2) * r12 + (double)g2 * r12;
            if (a.dtType == SNH_DT_DYNAMIC) im += mi;
        }
        if (a.dtType != SNH_DT_DYNAMIC) im = 1.0; /* host rejects other dt types in mixed mode */
        /* damping (or the prescribed damping branch), velocity, BC,
           coordinates: the FP64 node update */
        vmag = node_finish(a, n, f0, f1, f2, im, own, dt);
    }
    double bm = block_max(vmag, sh);
    if (threadIdx.x == 0) a.part_vmax[blk] = bm;
}

__global__ void __launch_bounds__(NODE_BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2))) k_node_mx(StepArgs a) {
    node_pass_mx(a, (int)blockIdx.x, (int)gridDim.x);
}

#endif
