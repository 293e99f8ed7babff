/*
 * snh_adjoint.h -- reverse-mode (adjoint) kernels of the explicit step,
 * FP64.  Included by snac_hip.hip after the forward kernels, whose helper
 * functions the forward recomputation below calls so that every branch
 * (yield, return branch, viscosity clamp) is taken exactly as in the forward
 * run.
 *
 * The map differentiated is the step with its discrete decisions frozen at
 * the recorded trajectory: dt, the rotation dt and the speed of sound of each
 * step, the damping branch sign(v) sign(f) of each node component
 * (UpdateNode.c:102-104), the plastic branch (none / shear / tension) of each
 * tet (plastic_BI/Constitutive.c:159-205) and the viscosity clamp
 * (maxwell/Constitutive.c:147-151).  Within those branches the derivative
 * is exact:
 *   - geometry uses n A = a / 2 with a the face cross product, so the
 *     velocity gradient of a tet is L = -G / (6 V), G_pq = sum_k u_kp a_(k+1)%4,q;
 *     strain rate = sym(L), Jaumann spin = skew(L) (StrainRate.c:65-184,
 *     UpdateElement.c:145-219);
 *   - the Mohr-Coulomb return F(S) = sum_k g_k(s) c_k c_k^T is differentiated
 *     in the eigenbasis: diagonal d g_k / d s_j, off-diagonal
 *     (g_k - g_l) / (s_k - s_l) (Daleckii-Krein);
 *   - the Maxwell update is linear in the stress and strain; its viscosity
 *     derivative is d sigma'/d ln(eta) = vic2 (sigma'_old + sigma'_new) temp
 *     with temp = mu dt / (2 eta).
 *
 * One reverse step n is two launches: k_node_adj (per node) then
 * k_elem_adj (one lane per tet, 6 elements per wave, as k_elem).
 */
#ifndef SNH_ADJOINT_H
#define SNH_ADJOINT_H

/* node slot (position of the element around the node, Force.c order) ->
   hex-local index of the node in that element */
static constexpr unsigned char kSlotHex[8] = {6, 7, 5, 4, 2, 3, 1, 0};

/* LDS slot of tet t's corner q in the per element-node sums: r * 8 + h for
   the hex node h = kTetraToNode[t][q] and the rank r of t among the 5 tets
   holding h (ascending).  Stored as [component][r][element][h], the 5 partial
   values of one element node sit 8 * EPW words apart, and lanes summing
   consecutive (element node, component) pairs fall in distinct LDS banks
   for the padded component strides below.  Packed 6 bits per tet, one word
   per corner. */
constexpr unsigned long long corner_slot_word(int q) {
    unsigned long long w = 0;
    for (int t = 0; t < NTET; t++) {
        const int h = kTetraToNode[t][q];
        int r = 0;
        for (int u = 0; u < t; u++)
            for (int p = 0; p < 4; p++)
                if (kTetraToNode[u][p] == h) r++;
        w |= (unsigned long long)(r * 8 + h) << (6 * t);
    }
    return w;
}
constexpr bool corner_slots_are_a_bijection() {
    bool seen[40] = {};
    for (int q = 0; q < 4; q++)
        for (int t = 0; t < NTET; t++) {
            const int s = (int)((corner_slot_word(q) >> (6 * t)) & 63ull);
            if (s >= 40 || seen[s]) return false;
            seen[s] = true;
        }
    return true;
}
static_assert(corner_slots_are_a_bijection(), "each hex node is held by 5 tets");
__device__ __forceinline__ int corner_slot(int t, int q) {
    constexpr unsigned long long W[4] = {corner_slot_word(0), corner_slot_word(1), corner_slot_word(2),
                                         corner_slot_word(3)};
    const int s = (int)((W[q] >> (6 * t)) & 63ull);
    return (s >> 3) * (EPW * 8) + (s & 7);
}

/* number of per element-node adjoint components written by k_elem_adj:
   coordinates (3), velocity (3), coordinates of the previous step (3) */
#define NBC 9
/* LDS words per component of the element-node sums in k_elem_adj:
   5 * EPW * 8 = 240 plus padding against bank conflicts.  The kernel's LDS
   (20,336 B per 64-lane block) must stay within 20 KiB for 8 blocks, two
   waves per SIMD, per CU. */
#define ADJ_LDS_STRIDE 249

struct AdjArgs {
    StepArgs f;           /* the forward arguments of the step reversed       */
    double* sbar;         /* [6][10][E] in: d/d sigma*_n; out: d/d sigma*_n-1 */
    double* psbar;        /* [10][E] in: d/d ps after step; out: before step  */
    const double* fbar;   /* [3][N] adjoint of the undamped nodal force        */
    const double* mbar;   /* [N]    adjoint of the inertial mass               */
    double* grad;         /* [E]    accumulated parameter gradient             */
    double* nb;           /* [E][8][NBC] per element-node contributions        */
    int param;            /* SNH_PARAM_*                                       */
    int flags;            /* SNH_ADJ_*                                         */
};

struct NodeAdjArgs {
    const double* nb;     /* [E][8][NBC] from the later step, or null          */
    double* xbar;         /* [3][N] in: d/d x_n+1 ; out: d/d x_n (direct part) */
    double* vbar;         /* [3][N] in: d/d v_n+1 ; out: d/d v_n (direct part) */
    double* fbar;         /* [3][N] out                                        */
    double* mbar;         /* [N]    out                                        */
    const double* tape_f;
    const double* tape_m;
    const uint8_t* tape_code;
    const uint8_t* bc;
    const DevBC* bcs;        /* velocity-condition table */
    const DevScalars* scal;  /* scalars of step n */
    double demf;
    int I, J, K, EI, EJ, EK, N, E;
};

/* a = (c2 - c1) x (c3 - c1): accumulate d/dc1, d/dc2, d/dc3 from d/da */
SNH_HD void face_cross_rev(const double* c1, const double* c2, const double* c3, const double ab[3], double* b1,
                           double* b2, double* b3) {
#pragma clang fp contract(fast)
    const double u0 = c2[0] - c1[0], u1 = c2[1] - c1[1], u2 = c2[2] - c1[2];
    const double w0 = c3[0] - c1[0], w1 = c3[1] - c1[1], w2 = c3[2] - c1[2];
    /* d/du = w x ab, d/dw = ab x u */
    const double ub0 = w1 * ab[2] - w2 * ab[1], ub1 = w2 * ab[0] - w0 * ab[2], ub2 = w0 * ab[1] - w1 * ab[0];
    const double wb0 = ab[1] * u2 - ab[2] * u1, wb1 = ab[2] * u0 - ab[0] * u2, wb2 = ab[0] * u1 - ab[1] * u0;
    b2[0] += ub0; b2[1] += ub1; b2[2] += ub2;
    b3[0] += wb0; b3[1] += wb1; b3[2] += wb2;
    b1[0] -= ub0 + wb0; b1[1] -= ub1 + wb1; b1[2] -= ub2 + wb2;
}

/* V = |(x3 - x0) . ((x1 - x0) x (x2 - x0))| / 6: accumulate d/dx from d/dV */
SNH_HD void tet_volume_rev(const double X[4][3], double Vb, double Xb[4][3]) {
#pragma clang fp contract(fast)
    double a0, a1, a2;
    face_cross(X[0], X[1], X[2], a0, a1, a2);
    const double d0 = X[3][0] - X[0][0], d1 = X[3][1] - X[0][1], d2 = X[3][2] - X[0][2];
    const double det = d0 * a0 + d1 * a1 + d2 * a2;
    const double g = (det < 0.0 ? -Vb : Vb) / 6.0;
    Xb[3][0] += g * a0; Xb[3][1] += g * a1; Xb[3][2] += g * a2;
    Xb[0][0] -= g * a0; Xb[0][1] -= g * a1; Xb[0][2] -= g * a2;
    const double ab[3] = {g * d0, g * d1, g * d2};
    face_cross_rev(X[0], X[1], X[2], ab, Xb[0], Xb[1], Xb[2]);
}

/* reverse of rotate_stress (UpdateElement.c:245-265): s is the stress before
   the rotation; b holds d/d(rotated stress) on entry and d/d(stress before)
   on exit; wb accumulates d/dw */
SNH_HD void rotate_stress_rev(const double s[6], const double w[3], double dt, double b[6], double wb[3]) {
#pragma clang fp contract(fast)
    const double w01 = w[0], w02 = w[1], w12 = w[2];
    /* forward values after each in-order update */
    const double f0 = s[0] + (2.0 * w01 * s[3] + 2.0 * w02 * s[4]) * dt;
    const double f1 = s[1] + (-2.0 * w01 * s[3] + 2.0 * w12 * s[5]) * dt;
    const double f2 = s[2] + (-2.0 * w02 * s[4] - 2.0 * w12 * s[5]) * dt;
    const double f3 = s[3] + w01 * (f1 - f0) * dt + (w02 * s[5] + w12 * s[4]) * dt;
    const double f4 = s[4] + w02 * (f2 - f0) * dt + (w01 * s[5] - w12 * f3) * dt;
    /* s5 += w12 (f2 - f1) dt + (-w01 f4 - w02 f3) dt */
    {
        const double g = b[5] * dt;
        b[2] += g * w12; b[1] -= g * w12; b[4] -= g * w01; b[3] -= g * w02;
        wb[2] += g * (f2 - f1); wb[0] -= g * f4; wb[1] -= g * f3;
    }
    /* s4 += w02 (f2 - f0) dt + (w01 s5 - w12 f3) dt */
    {
        const double g = b[4] * dt;
        b[2] += g * w02; b[0] -= g * w02; b[5] += g * w01; b[3] -= g * w12;
        wb[1] += g * (f2 - f0); wb[0] += g * s[5]; wb[2] -= g * f3;
    }
    /* s3 += w01 (f1 - f0) dt + (w02 s5 + w12 s4) dt */
    {
        const double g = b[3] * dt;
        b[1] += g * w01; b[0] -= g * w01; b[5] += g * w02; b[4] += g * w12;
        wb[0] += g * (f1 - f0); wb[1] += g * s[5]; wb[2] += g * s[4];
    }
    /* s2 += (-2 w02 s4 - 2 w12 s5) dt */
    {
        const double g = b[2] * dt;
        b[4] -= 2.0 * g * w02; b[5] -= 2.0 * g * w12;
        wb[1] -= 2.0 * g * s[4]; wb[2] -= 2.0 * g * s[5];
    }
    /* s1 += (-2 w01 s3 + 2 w12 s5) dt */
    {
        const double g = b[1] * dt;
        b[3] -= 2.0 * g * w01; b[5] += 2.0 * g * w12;
        wb[0] -= 2.0 * g * s[3]; wb[2] += 2.0 * g * s[5];
    }
    /* s0 += (2 w01 s3 + 2 w02 s4) dt */
    {
        const double g = b[0] * dt;
        b[3] += 2.0 * g * w01; b[4] += 2.0 * g * w02;
        wb[0] += 2.0 * g * s[3]; wb[1] += 2.0 * g * s[4];
    }
}

/* reverse of the tet kinematics, first half: from d/d(strain rate) srb
   (00,11,22,01,02,12) and d/d(spin) wb (01,02,12), the adjoint of the
   face-vector sums G (L = -G / (6V), so Gb = -Lb / (6V)) and the direct
   volume term Vb -= Lb:L / V */
SNH_HD void tet_kinematics_rev_l(const TetKin& k, const double srb[6], const double wb[3], double& Vb,
                                 double Gb[3][3]) {
#pragma clang fp contract(fast)
    double L[3][3], Lb[3][3];
    L[0][0] = k.sr[0]; L[1][1] = k.sr[1]; L[2][2] = k.sr[2];
    L[0][1] = k.sr[3] + k.w[0]; L[1][0] = k.sr[3] - k.w[0];
    L[0][2] = k.sr[4] + k.w[1]; L[2][0] = k.sr[4] - k.w[1];
    L[1][2] = k.sr[5] + k.w[2]; L[2][1] = k.sr[5] - k.w[2];
    Lb[0][0] = srb[0]; Lb[1][1] = srb[1]; Lb[2][2] = srb[2];
    Lb[0][1] = 0.5 * (srb[3] + wb[0]); Lb[1][0] = 0.5 * (srb[3] - wb[0]);
    Lb[0][2] = 0.5 * (srb[4] + wb[1]); Lb[2][0] = 0.5 * (srb[4] - wb[1]);
    Lb[1][2] = 0.5 * (srb[5] + wb[2]); Lb[2][1] = 0.5 * (srb[5] - wb[2]);
    const double V = k.V;
    double LL = 0.0;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) LL += Lb[p][q] * L[p][q];
    Vb -= LL / V;
    const double gs = -1.0 / (6.0 * V);
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) Gb[p][q] = gs * Lb[p][q];
}

/* second half: from Gb, the volume adjoint Vb and d/d(face cross products)
   ab, accumulate d/dX and d/dU of the 4 nodes (G_pq = sum_k u_kp a_(k+1)%4,q,
   V = |(x3 - x0) . a0| / 6) */
SNH_HD void tet_kinematics_rev_x(const double X[4][3], const double U[4][3], const double Gb[3][3], double Vb,
                                 double ab[4][3], double Xb[4][3], double Ub[4][3]) {
#pragma clang fp contract(fast)
    double a[4][3];
#pragma unroll
    for (int f = 0; f < 4; f++)
        face_cross(X[kFaceToNode[f][0]], X[kFaceToNode[f][1]], X[kFaceToNode[f][2]], a[f][0], a[f][1], a[f][2]);
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
        const int f = (kk + 1) & 3;
#pragma unroll
        for (int p = 0; p < 3; p++) {
            double ub = 0.0;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                ub += Gb[p][q] * a[f][q];
                ab[f][q] += Gb[p][q] * U[kk][p];
            }
            Ub[kk][p] += ub;
        }
    }
    const double d0 = X[3][0] - X[0][0], d1 = X[3][1] - X[0][1], d2 = X[3][2] - X[0][2];
    const double det = d0 * a[0][0] + d1 * a[0][1] + d2 * a[0][2];
    const double g = (det < 0.0 ? -Vb : Vb) / 6.0;
    Xb[3][0] += g * a[0][0]; Xb[3][1] += g * a[0][1]; Xb[3][2] += g * a[0][2];
    Xb[0][0] -= g * a[0][0]; Xb[0][1] -= g * a[0][1]; Xb[0][2] -= g * a[0][2];
    ab[0][0] += g * d0; ab[0][1] += g * d1; ab[0][2] += g * d2;
#pragma unroll
    for (int f = 0; f < 4; f++)
        face_cross_rev(X[kFaceToNode[f][0]], X[kFaceToNode[f][1]], X[kFaceToNode[f][2]], ab[f], Xb[kFaceToNode[f][0]],
                       Xb[kFaceToNode[f][1]], Xb[kFaceToNode[f][2]]);
}

/* the branch mc_return takes, with its multiplier: 0 elastic, 1 shear,
   2 tension (same expressions, same order as mc_return) */
struct MCBranch {
    int branch;
    double alam, anphi, anpsi, den;
};
SNH_HD MCBranch mc_branch(const DevMaterial& m, const double s[3], double sphi, double spsi, double cohesion,
                          double tension_cutoff) {
    MCBranch r;
    const double a1 = m.lambda + 2.0 * m.mu;
    const double a2 = m.lambda;
    r.anphi = (1.0 + sphi) / (1.0 - sphi);
    r.anpsi = (1.0 + spsi) / (1.0 - spsi);
    r.den = a1 - a2 * r.anpsi + a1 * r.anphi * r.anpsi - a2 * r.anphi;
    r.alam = 0.0;
    r.branch = 0;
    const double fs = s[0] - s[2] * r.anphi + 2 * cohesion * sqrt(r.anphi);
    const double ft = s[2] - tension_cutoff;
    if (!(fs < 0.0 || ft > 0.0)) return r;
    const double aP = sqrt(1.0 + r.anphi * r.anphi) + r.anphi;
    const double sP = tension_cutoff * r.anphi - 2 * cohesion * sqrt(r.anphi);
    const double h = s[2] - tension_cutoff + aP * (s[0] - sP);
    if (h < 0.0) {
        r.branch = 1;
        r.alam = fs / r.den;
    } else {
        r.branch = 2;
        r.alam = ft / a1;
    }
    return r;
}

/* reverse of the Mohr-Coulomb return and back-rotation.  s, cn: eigenvalues
   (ascending) and eigenvectors of the trial stress; ob: d/d(returned stress,
   6 components); psb: d/d(plastic strain after the step).  Writes d/d(trial
   stress) to Sb and accumulates d/d(cohesion) and d/d(tension cut-off). */
SNH_HD void mc_return_rev(const DevMaterial& m, const MCBranch& br, const double s[3], const double cn[3][3],
                          const double ob[6], double psb, double Sb[6], double& cohb, double& tcb, bool clamp) {
#pragma clang fp contract(fast)
    const double a1 = m.lambda + 2.0 * m.mu;
    const double a2 = m.lambda;
    double beta[3], K;
    if (br.branch == 1) {
        beta[0] = a1 - a2 * br.anpsi;
        beta[1] = a2 * (1.0 - br.anpsi);
        beta[2] = a2 - a1 * br.anpsi;
        const double mm = (1.0 - br.anpsi) / 3.0;
        K = sqrt(0.5 * ((1.0 - mm) * (1.0 - mm) + mm * mm + (br.anpsi + mm) * (br.anpsi + mm) + mm * mm));
    } else {
        beta[0] = a2;
        beta[1] = a2;
        beta[2] = a1;
        K = sqrt(0.5 * 7.0 / 9.0);
    }
    /* returned stress adjoint in the eigenbasis: Mb = C^T Ob C */
    double Ob[3][3];
    Ob[0][0] = ob[0]; Ob[1][1] = ob[1]; Ob[2][2] = ob[2];
    Ob[0][1] = Ob[1][0] = 0.5 * ob[3];
    Ob[0][2] = Ob[2][0] = 0.5 * ob[4];
    Ob[1][2] = Ob[2][1] = 0.5 * ob[5];
    double T[3][3], Mb[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int l = 0; l < 3; l++) T[i][l] = Ob[i][0] * cn[0][l] + Ob[i][1] * cn[1][l] + Ob[i][2] * cn[2][l];
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
        for (int l = 0; l < 3; l++) Mb[kk][l] = cn[0][kk] * T[0][l] + cn[1][kk] * T[1][l] + cn[2][kk] * T[2][l];
    /* g_k = s_k - alam beta_k ; plastic strain += K |alam| */
    double alamb = -(Mb[0][0] * beta[0] + Mb[1][1] * beta[1] + Mb[2][2] * beta[2]);
    alamb += psb * K * (br.alam < 0.0 ? -1.0 : 1.0);
    double Db[3][3];
    Db[0][0] = Mb[0][0]; Db[1][1] = Mb[1][1]; Db[2][2] = Mb[2][2];
    if (br.branch == 1) {
        /* alam = (s0 - s2 anphi + 2 c sqrt(anphi)) / den */
        Db[0][0] += alamb / br.den;
        Db[2][2] -= alamb * br.anphi / br.den;
        cohb += alamb * 2.0 * sqrt(br.anphi) / br.den;
    } else {
        /* alam = (s2 - tc) / a1 */
        Db[2][2] += alamb / a1;
        tcb -= alamb / a1;
    }
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
#pragma unroll
        for (int l = kk + 1; l < 3; l++) {
            const double gap = s[kk] - s[l];
            double G = 1.0;
            if (fabs(gap) > 1e-14 * (fabs(s[kk]) + fabs(s[l]))) G = 1.0 - br.alam * (beta[kk] - beta[l]) / gap;
            if (clamp) G = G < 0.0 ? 0.0 : (G > 1.0 ? 1.0 : G);
            Db[kk][l] = Db[l][kk] = G * Mb[kk][l];
        }
    /* Sb = C Db C^T */
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int l = 0; l < 3; l++) T[i][l] = cn[i][0] * Db[0][l] + cn[i][1] * Db[1][l] + cn[i][2] * Db[2][l];
    double S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] = T[i][0] * cn[j][0] + T[i][1] * cn[j][1] + T[i][2] * cn[j][2];
    Sb[0] = S[0][0];
    Sb[1] = S[1][1];
    Sb[2] = S[2][2];
    Sb[3] = S[0][1] + S[1][0];
    Sb[4] = S[0][2] + S[2][0];
    Sb[5] = S[1][2] + S[2][1];
}

/* ------------------------------------------------------------------------
 * k_node_adj: reverse of the node pass of step n (UpdateNode.c:60-158 and
 * the velocity/coordinate update), plus the gather of the element-node
 * contributions of step n+1 in slot order.
 * ---------------------------------------------------------------------- */
__global__ void __launch_bounds__(NODE_BLOCK) k_node_adj(NodeAdjArgs A) {
    const int n = xcd_block() * blockDim.x + threadIdx.x;
    if (n >= A.N) return;
    const size_t N = (size_t)A.N;
    double xb[3] = {A.xbar[n], A.xbar[N + n], A.xbar[2 * N + n]};
    double vb[3] = {A.vbar[n], A.vbar[N + n], A.vbar[2 * N + n]};
    double xpb[3] = {0.0, 0.0, 0.0};
    if (A.nb) {
        const int i = n % A.I, j = (n / A.I) % A.J, k = n / (A.I * A.J);
#pragma unroll
        for (int slot = 0; slot < 8; slot++) {
            const int ei = i - 1 + (slot & 1), ej = j - 1 + ((slot >> 1) & 1), ek = k - 1 + ((slot >> 2) & 1);
            if (ei < 0 || ej < 0 || ek < 0 || ei >= A.EI || ej >= A.EJ || ek >= A.EK) continue;
            const size_t e = (size_t)ei + (size_t)A.EI * (ej + (size_t)A.EJ * ek);
            const double* p = A.nb + (e * 8 + kSlotHex[slot]) * NBC;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                xb[d] += p[d];
                vb[d] += p[3 + d];
                xpb[d] += p[6 + d];
            }
        }
    }
    const double dt = A.scal->dt;
    /* x_n+1 = x_n + dt v_n+1 ; x_n is also the previous coordinate of step n+1 */
    double vp[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        A.xbar[d * N + n] = xb[d] + xpb[d];
        vp[d] = vb[d] + dt * xb[d];
    }
    const uint32_t code = A.bc[n];
    if (code) {
        const DevBC& b = A.bcs[code];
        if (b.mask & 1u) vp[0] = 0.0;
        if (b.mask & 2u) vp[1] = 0.0;
        if (b.mask & 4u) vp[2] = 0.0;
    }
    /* v_n+1 = v_n + dt k f / m, k = 1 - demf sign(v) sign(f) */
    const double m = A.tape_m[n];
    const uint8_t tc = A.tape_code[n];
    double mb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double kd = 1.0 - A.demf * (double)((int)((tc >> (2 * d)) & 3u) - 1);
        const double g = dt * kd * vp[d] / m;
        A.vbar[d * N + n] = vp[d];
        A.fbar[d * N + n] = g;
        mb -= g * A.tape_f[d * N + n] / m;
    }
    A.mbar[n] = mb;
}

/* ------------------------------------------------------------------------
 * k_elem_adj: reverse of the element pass of step n (pending rotation,
 * strain rate, constitutive update, trace averaging) and of the element
 * terms of the node pass (force sigma.a/12, gravity, inertial mass;
 * Force.c:112-173).  It recomputes the forward element pass, then reverses
 * it in two phases that share few registers: the constitutive phase holds
 * the kinematics and the stresses, the geometry phase loads the node
 * coordinates and velocities again.  Each phase fits 256 VGPRs, so the
 * kernel runs two waves per SIMD; the bound below holds the two-law
 * variants there too (elastic with plastic: 256 VGPRs, no spill; all three
 * laws: 144 B of spills per lane instead of one wave per SIMD).
 * ---------------------------------------------------------------------- */
template <int RSET>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) k_elem_adj(AdjArgs A) {
    const StepArgs& a = A.f;
    __shared__ double sh_a[EPW][NTET];
    __shared__ double sh_b[EPW][NTET];
    __shared__ double sh_sp[EPW][NTET][3];   /* matched, own cohesion, slope */
    __shared__ double sh_nb[NBC * ADJ_LDS_STRIDE];   /* [comp][r][element][h] */
    const int lane = threadIdx.x;
    const bool in_wave = lane < EPW * NTET;
    const int le = in_wave ? lane / NTET : 0;
    const int t = in_wave ? lane - le * NTET : 0;
    const int lb = xcd_block();
    const int e = lb * EPW + le;
    const bool active = in_wave && e < a.E;
    const size_t E = (size_t)a.E;
    const int ec = active ? e : 0;
    int ei, ej, ek;
    elem_ijk(a, ec, ei, ej, ek);
    const double dt = a.scal->dt;
    const double dt_rot = a.scal->dt_rot;
    const double csound = a.scal->speedOfSound;
    const DevMaterial& m = a.mats[a.mat ? a.mat[ec] : 0];

    /* ---------------- forward recomputation (as k_elem) ---------------- */
    int nd[4];
    double Xt[4][3], Ut[4][3];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        nd[q] = hex_node_index(a, ei, ej, ek, tet_node(t, q));
        Xt[q][0] = a.xc[nd[q]];
        Xt[q][1] = a.xc[a.N + nd[q]];
        Xt[q][2] = a.xc[2 * a.N + nd[q]];
        Ut[q][0] = a.v[nd[q]];
        Ut[q][1] = a.v[a.N + nd[q]];
        Ut[q][2] = a.v[2 * a.N + nd[q]];
    }
    TetKin k;
    tet_kinematics(Xt, Ut, k);
    /* node pass terms of the stress adjoint: f += sigma a_f / 12 at the node
       opposite face f, f_y -= g rho V / 8, m += (lambda + 2 mu) / c^2 V / 4
       (Force.c:112-173); taken here, while the coordinates are live */
    double ob[6], Vb = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) ob[c] = active ? A.sbar[((size_t)c * NTET + t) * E + ec] : 0.0;
    {
#pragma clang fp contract(fast)
        double fa[4][3];
#pragma unroll
        for (int f = 0; f < 4; f++)
            face_cross(Xt[kFaceToNode[f][0]], Xt[kFaceToNode[f][1]], Xt[kFaceToNode[f][2]], fa[f][0], fa[f][1],
                       fa[f][2]);
        const double gr = a.gravity * m.rho_eff / 8.0;
        const double md = 0.25 * (m.lam2mu / (csound * csound));
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int f = (q + 1) & 3;
            const double b0 = A.fbar[nd[q]], b1 = A.fbar[a.N + nd[q]], b2 = A.fbar[2 * a.N + nd[q]];
            const double mb = A.mbar[nd[q]];
            const double c12 = 1.0 / 12.0;
            ob[0] += c12 * b0 * fa[f][0];
            ob[1] += c12 * b1 * fa[f][1];
            ob[2] += c12 * b2 * fa[f][2];
            ob[3] += c12 * (b0 * fa[f][1] + b1 * fa[f][0]);
            ob[4] += c12 * (b0 * fa[f][2] + b2 * fa[f][0]);
            ob[5] += c12 * (b1 * fa[f][2] + b2 * fa[f][1]);
            Vb += -gr * b1 + md * mb;
        }
    }
    const double trace = snh_div3(k.sr[0] + k.sr[1] + k.sr[2]);
    if (in_wave) {
        sh_a[le][t] = k.V;
        sh_b[le][t] = trace * k.V;
    }
    __syncthreads();
    double evol = 0.0, partial = 0.0;
#pragma unroll
    for (int q = 0; q < NTET; q++) {
        evol += sh_a[le][q];
        partial += sh_b[le][q];
    }
    evol *= 0.5;
    const double elemVolume = 2.0 * evol;
    const double srcorr = partial / elemVolume;

    double st_in[6], st[6];
#pragma unroll
    for (int c = 0; c < 6; c++) st_in[c] = active ? a.stress[((size_t)c * NTET + t) * E + ec] : 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) st[c] = st_in[c];
    rotate_stress(st, k.w, dt_rot);
    double st_rot[6];
#pragma unroll
    for (int c = 0; c < 6; c++) st_rot[c] = st[c];
    double sn[6];
#pragma unroll
    for (int c = 0; c < 3; c++) sn[c] = (k.sr[c] + (-1.0 * trace + srcorr)) * dt;
#pragma unroll
    for (int c = 3; c < 6; c++) sn[c] = k.sr[c] * dt;

    const bool is_pl = (RSET & RS_PLASTIC) && (m.rheology & SNH_RHEOLOGY_PLASTIC);
    const bool is_mx = !is_pl && (RSET & RS_MAXWELL) && (m.rheology & SNH_RHEOLOGY_MAXWELL);
    /* plastic: softened cohesion and where it came from */
    double ps = 0.0, own_slope = 0.0, coh = 0.0, tc = 0.0, sphi = 0.0, spsi = 0.0, phi = 0.0;
    int src = -1;
    MCBranch br = {0, 0.0, 0.0, 0.0, 0.0};
    double d3[3] = {0.0, 0.0, 0.0}, cn[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    if (RSET & RS_PLASTIC) {
        double matched = 0.0, own_coh = 0.0, own_phi = 0.0, own_psi = 0.0;
        if (is_pl) {
            if (active) ps = a.ps[(size_t)t * E + ec];
            for (uint32_t i = 0; i < m.nseg; i++) {
                const double pl1 = m.plstrain[i], pl2 = m.plstrain[i + 1];
                if (ps >= pl1 && ps <= pl2) {
                    const double tgf = m.tgf[i];
                    const double tgd = m.tgd[i];
                    const double tgc = m.tgc[i];
                    own_phi = m.fa[i] + tgf * (ps - pl1);
                    own_psi = m.da[i] + tgd * (ps - pl1);
                    own_coh = m.coh[i] + tgc * (ps - pl1);
                    own_slope = tgc;
                    matched = 1.0;
                }
            }
        }
        __syncthreads();   /* sh_a still held the tet volumes */
        if (in_wave) {
            sh_sp[le][t][0] = matched;
            sh_sp[le][t][1] = own_coh;
            sh_sp[le][t][2] = own_phi;
            sh_a[le][t] = own_psi;
        }
        __syncthreads();
        if (is_pl) {
            double psi = 0.0, pc = 30.0e6;
            phi = 30.0;
            for (int q = 0; q <= t; q++)
                if (sh_sp[le][q][0] != 0.0) {
                    phi = sh_sp[le][q][2];
                    psi = sh_a[le][q];
                    pc = sh_sp[le][q][1];
                    src = q;
                }
            coh = a.cohoff ? pc + a.cohoff[ec] : pc;
            tc = m.ten_off;
            if (phi > 0.0 && tc == 0.0) tc = coh / tan_phi(m, phi);
            angle_terms(m, phi, psi, sphi, spsi);
            double trace_strain = sn[0] + sn[1] + sn[2];
            st[0] += (2.0 * m.mu) * sn[0] + m.lambda * (trace_strain);
            st[1] += (2.0 * m.mu) * sn[1] + m.lambda * (trace_strain);
            st[2] += (2.0 * m.mu) * sn[2] + m.lambda * (trace_strain);
            st[3] += (2.0 * m.mu) * sn[3];
            st[4] += (2.0 * m.mu) * sn[4];
            st[5] += (2.0 * m.mu) * sn[5];
            double S[3][3];
            S[0][0] = st[0]; S[1][1] = st[1]; S[2][2] = st[2];
            S[0][1] = S[1][0] = st[3];
            S[0][2] = S[2][0] = st[4];
            S[1][2] = S[2][1] = st[5];
            if (!mc_surely_elastic(m, st, sphi, coh, tc)) {
                principal_stresses(S, d3, cn);
                br = mc_branch(m, d3, sphi, spsi, coh, tc);
                double dps = 0.0;
                double s3[3] = {d3[0], d3[1], d3[2]};
                if (mc_return(m, S, s3, cn, sphi, spsi, coh, tc, 0.0, false, dps)) {
                    st[0] = S[0][0]; st[1] = S[1][1]; st[2] = S[2][2];
                    st[3] = S[0][1]; st[4] = S[0][2]; st[5] = S[1][2];
                }
            }
        }
    }
    /* Maxwell */
    double Xp[4][3], te = 0.0, old_volume = 1.0, visc = 1.0;
    bool clamped = true;
    if (is_mx) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Xp[q][0] = a.xp[nd[q]];
            Xp[q][1] = a.xp[a.N + nd[q]];
            Xp[q][2] = a.xp[2 * a.N + nd[q]];
        }
        old_volume = tet_volume(Xp[0], Xp[1], Xp[2], Xp[3]);
        te = k.V / old_volume - 1.0;
        double d0, d1, d2;
        devstrain(sn, te, d0, d1, d2);
        const double rvisc = a.refvisc ? a.refvisc[ec] : m.refvisc;
        visc = viscosity_of(m, rvisc, d0, d1, d2, sn, dt);
        clamped = !(rvisc >= m.vis_min && rvisc <= m.vis_max);
        maxwell_update(m, st, sn, te, visc, dt);
    } else if (!is_pl && (RSET & RS_ELASTIC) && (m.rheology & SNH_RHEOLOGY_ELASTIC)) {
        const double a2 = m.lambda;
        const double sh2 = 2.0 * m.mu;
        double trace_strain = sn[0] + sn[1] + sn[2];
        st[0] += sh2 * sn[0] + a2 * (trace_strain);
        st[1] += sh2 * sn[1] + a2 * (trace_strain);
        st[2] += sh2 * sn[2] + a2 * (trace_strain);
        st[3] += sh2 * sn[3];
        st[4] += sh2 * sn[4];
        st[5] += sh2 * sn[5];
    }
    /* trace averaging */
    const double tr = snh_div3(st[0] + st[1] + st[2]);
    __syncthreads();
    if (in_wave) sh_b[le][t] = tr * k.V;
    __syncthreads();
    double pst = 0.0;
#pragma unroll
    for (int q = 0; q < NTET; q++) pst += sh_b[le][q];
    const double avgTrace = pst / elemVolume;

    /* ---------------- reverse ----------------
       (arithmetic of the adjoint only: the forward values above are exact
       recomputations and keep the uncontracted rounding of the forward) */
    {
#pragma clang fp contract(fast)
        /* trace averaging: out_ii = s_ii - tr + pst / EV */
        const double dsum = ob[0] + ob[1] + ob[2];
        __syncthreads();
        if (in_wave) sh_b[le][t] = active ? dsum : 0.0;
        __syncthreads();
        double avgb = 0.0;
#pragma unroll
        for (int q = 0; q < NTET; q++) avgb += sh_b[le][q];
        const double pstb = avgb / elemVolume;
        double EVb = -avgb * avgTrace / elemVolume;
        double trb = -dsum + pstb * k.V;
        Vb += pstb * tr;
        double sg[6];
#pragma unroll
        for (int c = 0; c < 6; c++) sg[c] = ob[c];
        sg[0] += trb / 3.0;
        sg[1] += trb / 3.0;
        sg[2] += trb / 3.0;

        /* constitutive: d/d(rotated stress) strb, d/d(strain increment) snb */
        double strb[6], snb[6], parb = 0.0, psb_self = 0.0, cohb = 0.0;
        if (is_pl) {
            double psb = active ? A.psbar[(size_t)t * E + ec] : 0.0;
            psb_self = psb;
            double Sb[6];
            if (br.branch) {
                double tcb = 0.0;
                mc_return_rev(m, br, d3, cn, sg, psb, Sb, cohb, tcb, (A.flags & SNH_ADJ_CORNER_CLAMP) != 0);
                if (m.ten_off == 0.0 && phi > 0.0) cohb += tcb / tan_phi(m, phi);
            } else {
#pragma unroll
                for (int c = 0; c < 6; c++) Sb[c] = sg[c];
            }
            const double trS = Sb[0] + Sb[1] + Sb[2];
#pragma unroll
            for (int c = 0; c < 6; c++) {
                strb[c] = Sb[c];
                snb[c] = 2.0 * m.mu * Sb[c];
            }
            snb[0] += m.lambda * trS;
            snb[1] += m.lambda * trS;
            snb[2] += m.lambda * trS;
            if (A.param == SNH_PARAM_COHESION) parb = cohb;
        } else if (is_mx) {
            const double rmu = m.mu;
            const double bulkm = m.lambda + 2.0 * m.mu / 3.0;
            const double temp = rmu / (2.0 * visc) * dt;
            const double vic1 = 1.0 - temp, vic2 = 1.0 / (1.0 + temp);
            const double ts3 = (st_rot[0] + st_rot[1] + st_rot[2]) / 3.0;
            const double te3 = te / 3.0;
            const double VSb = sg[0] + sg[1] + sg[2];
            double sdb_sum = 0.0, edb_sum = 0.0, thb = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double sd = st_rot[i] - ts3;
                const double outd = (sd * vic1 + 2.0 * rmu * (sn[i] - te3)) * vic2;
                const double sdb = sg[i] * vic1 * vic2;
                const double edb = sg[i] * 2.0 * rmu * vic2;
                strb[i] = sdb;
                snb[i] = edb;
                sdb_sum += sdb;
                edb_sum += edb;
                thb += sg[i] * vic2 * (sd + outd) * temp;
            }
#pragma unroll
            for (int i = 0; i < 3; i++) strb[i] += -sdb_sum / 3.0 + VSb / 3.0;
#pragma unroll
            for (int c = 3; c < 6; c++) {
                const double outc = (st_rot[c] * vic1 + 2.0 * rmu * sn[c]) * vic2;
                strb[c] = sg[c] * vic1 * vic2;
                snb[c] = sg[c] * 2.0 * rmu * vic2;
                thb += sg[c] * vic2 * (st_rot[c] + outc) * temp;
            }
            const double teb = -edb_sum / 3.0 + bulkm * VSb;
            /* te = V / V_old - 1 */
            Vb += teb / old_volume;
            const double Voldb = -teb * k.V / (old_volume * old_volume);
            double Xpb[4][3] = {{0.0}};
            tet_volume_rev(Xp, Voldb, Xpb);
            if (in_wave) {
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int d = 0; d < 3; d++) sh_nb[(6 + d) * ADJ_LDS_STRIDE + corner_slot(t, q) + le * 8] = Xpb[q][d];
            }
            if (A.param == SNH_PARAM_LOG_VISCOSITY && !clamped) parb = thb;
        } else {
            const double trS = sg[0] + sg[1] + sg[2];
#pragma unroll
            for (int c = 0; c < 6; c++) {
                strb[c] = sg[c];
                snb[c] = 2.0 * m.mu * sg[c];
            }
            snb[0] += m.lambda * trS;
            snb[1] += m.lambda * trS;
            snb[2] += m.lambda * trS;
        }
        if ((RSET & RS_MAXWELL) && !is_mx && in_wave) {
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int d = 0; d < 3; d++) sh_nb[(6 + d) * ADJ_LDS_STRIDE + corner_slot(t, q) + le * 8] = 0.0;
        }

        /* strain: sn = (sr + (-trace + srcorr) on the diagonal) dt */
        double srb[6];
#pragma unroll
        for (int c = 0; c < 6; c++) srb[c] = snb[c] * dt;
        const double d6 = srb[0] + srb[1] + srb[2];
        __syncthreads();
        if (in_wave) {
            sh_b[le][t] = active ? d6 : 0.0;
            sh_a[le][t] = active ? parb : 0.0;
            sh_sp[le][t][1] = active ? cohb : 0.0;
            sh_sp[le][t][2] = (double)src;
        }
        __syncthreads();
        double srcorrb = 0.0, gsum = 0.0;
#pragma unroll
        for (int q = 0; q < NTET; q++) {
            srcorrb += sh_b[le][q];
            gsum += sh_a[le][q];
        }
        const double partialb = srcorrb / elemVolume;
        EVb += -srcorrb * srcorr / elemVolume;
        const double traceb = -d6 + partialb * k.V;
        Vb += partialb * trace;
        srb[0] += traceb / 3.0;
        srb[1] += traceb / 3.0;
        srb[2] += traceb / 3.0;
        Vb += EVb;   /* elemVolume = sum of the tet volumes */

        /* plastic strain before the step: the identity (ps += dps) plus the
           softening slope for every tet whose cohesion came from this tet */
        if (is_pl && active) {
            double cs = 0.0;
            for (int q = t; q < NTET; q++)
                if (sh_sp[le][q][2] == (double)t) cs += sh_sp[le][q][1];
            A.psbar[(size_t)t * E + ec] = psb_self + own_slope * cs;
        }
        if (active && t == 0) A.grad[ec] += gsum;

        /* rotation of the stored stress */
        double wbv[3] = {0.0, 0.0, 0.0};
        double stb[6];
#pragma unroll
        for (int c = 0; c < 6; c++) stb[c] = strb[c];
        rotate_stress_rev(st_in, k.w, dt_rot, stb, wbv);
        if (active) {
#pragma unroll
            for (int c = 0; c < 6; c++) A.sbar[((size_t)c * NTET + t) * E + ec] = stb[c];
        }

        /* geometry: the coordinates and velocities are loaded again here
           rather than held across the constitutive reverse (the empty asm
           keeps the compiler from reusing the first loads), and sigma*_n is
           read from the replay tape */
        double Gb[3][3];
        tet_kinematics_rev_l(k, srb, wbv, Vb, Gb);
        int nr[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            nr[q] = nd[q];
            asm volatile("" : "+v"(nr[q]));
        }
        double Xr[4][3], Ur[4][3], ab[4][3] = {{0.0}};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Xr[q][0] = a.xc[nr[q]];
            Xr[q][1] = a.xc[a.N + nr[q]];
            Xr[q][2] = a.xc[2 * a.N + nr[q]];
            Ur[q][0] = a.v[nr[q]];
            Ur[q][1] = a.v[a.N + nr[q]];
            Ur[q][2] = a.v[2 * a.N + nr[q]];
        }
        {
            double so[6];   /* sigma*_n of this tet */
#pragma unroll
            for (int c = 0; c < 6; c++) so[c] = active ? a.stress_out[((size_t)c * NTET + t) * E + ec] : 0.0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int f = (q + 1) & 3;
                const double b0 = A.fbar[nr[q]], b1 = A.fbar[a.N + nr[q]], b2 = A.fbar[2 * a.N + nr[q]];
                const double c12 = 1.0 / 12.0;
                ab[f][0] += c12 * (so[0] * b0 + so[3] * b1 + so[4] * b2);
                ab[f][1] += c12 * (so[3] * b0 + so[1] * b1 + so[5] * b2);
                ab[f][2] += c12 * (so[4] * b0 + so[5] * b1 + so[2] * b2);
            }
        }
        double Xb[4][3] = {{0.0}}, Ub[4][3] = {{0.0}};
        tet_kinematics_rev_x(Xr, Ur, Gb, Vb, ab, Xb, Ub);
        if (in_wave) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int sl = corner_slot(t, q) + le * 8;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    sh_nb[d * ADJ_LDS_STRIDE + sl] = Xb[q][d];
                    sh_nb[(3 + d) * ADJ_LDS_STRIDE + sl] = Ub[q][d];
                }
            }
        }
        __syncthreads();
        /* per element-node sums over the 5 tets at each hex node, ascending
           tet order, into [E][8][NBC]; components 6..8 (coordinates of the
           step before) only where a Maxwell law is compiled in, else they
           stay zero */
        constexpr int NC = (RSET & RS_MAXWELL) ? NBC : 6;
        for (int o = lane; o < NC * EPW * 8; o += 64) {
            const int rem = o / NC;                 /* element * 8 + h */
            const int comp = o - rem * NC;
            if (lb * EPW + (rem >> 3) >= a.E) continue;
            double sum = 0.0;
#pragma unroll
            for (int r = 0; r < 5; r++) sum += sh_nb[comp * ADJ_LDS_STRIDE + r * (EPW * 8) + rem];
            A.nb[((size_t)lb * (EPW * 8) + rem) * NBC + comp] = sum;
        }
    }
}

#endif
