/*
 * snh_kernels.h -- device-side data structures and per-tetrahedron arithmetic
 * of the explicit step for gfx950 (FP64 path).
 *
 * Layout in HBM (structure of arrays, one thread per element or node):
 *   coordinates   [3][N]      two buffers (current / previous step)
 *   velocity      [3][N]
 *   stress        [6][10][E]  components 00,11,22,01,02,12 of each tetrahedron
 *   plastic strain[10][E]
 * Every expression below keeps the operand order of the reference line it
 * cites; the file is compiled with -ffp-contract=off.
 */
#ifndef SNH_KERNELS_H
#define SNH_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "snh_math.h"

namespace snh {

constexpr int NTET = 10;

/* Snac/libSnac/src/TetrahedraTables.c:42-84 (the reference's 2 x 5-tet
   decomposition of the hex; identical tables keep the stored tet stresses
   meaning the same). */
static constexpr unsigned char kTetraToNode[NTET][4] = {
    {7, 2, 3, 0}, {0, 5, 4, 7}, {2, 5, 1, 0}, {5, 2, 6, 7}, {2, 5, 0, 7},
    {3, 4, 7, 6}, {4, 3, 0, 1}, {6, 1, 2, 3}, {1, 6, 5, 4}, {1, 3, 6, 4}};
static constexpr unsigned char kFaceToNode[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {3, 1, 0}};
static constexpr unsigned char kNodeToTetra[8][5] = {
    {3, 5, 7, 8, 9}, {0, 1, 3, 4, 5}, {1, 2, 3, 4, 8}, {1, 5, 6, 8, 9},
    {0, 2, 3, 4, 7}, {0, 5, 6, 7, 9}, {2, 6, 7, 8, 9}, {0, 1, 2, 4, 6}};
static constexpr unsigned char kNodeToSurface[8][5] = {
    {3, 0, 1, 2, 3}, {1, 0, 0, 0, 3}, {2, 2, 1, 2, 3}, {3, 2, 1, 0, 0},
    {2, 1, 2, 1, 3}, {3, 1, 2, 0, 2}, {3, 0, 2, 1, 1}, {0, 1, 0, 3, 3}};
/* hex-local node -> (di, dj, dk) offset (HexaEL.c:462-469) */
static constexpr unsigned char kHexOff[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* double of the reference's float literals */
constexpr double F_045 = (double)0.45f;    /* Context.c:1251 */
constexpr double F_3EM11 = (double)3.0e-11f; /* Context.c:1219 */
constexpr double F_1EP27 = (double)1.0e+27f; /* UpdateElement.c:50 */

struct DevMaterial {
    uint32_t rheology;
    double lambda, mu;
    double rho_eff;          /* density after Stress.c:169-170 with beta = 0 */
    double lam2mu;           /* lambda + 2.0f*mu (Force.c:96) */
    /* plasticity */
    uint32_t nseg;
    double plstrain[9], fa[9], da[9], coh[9];
    /* slopes of each softening segment, (x[i+1] - x[i]) / (pl[i+1] - pl[i]),
       computed once on the host (the same correctly rounded quotients) */
    double tgf[8], tgd[8], tgc[8];
    double ten_off;
    double sphi, spsi, tanphi;  /* host libm sin/sin/tan of fa[0], da[0] */
    double anphi0, sqanphi0;    /* (1 + sphi) / (1 - sphi) and its root, for the
                                   elastic bound only */
    /* viscosity */
    double vis_min, vis_max, refvisc, refsrate, reftemp, activationE, srexponent;
    int linear_visc;         /* srexponent == 1 and activationE == 0 */
    /* temperature plugin; beta: pressure-dependent density (Stress.c:169-170) */
    double phsDensity, alpha, kcond, cp, beta;
};

struct DevScalars {
    double dt;             /* dt of the step being run                    */
    double dt_rot;         /* dt of the Jaumann rotation still to apply   */
    double speedOfSound;
    double minLen;
    double vmax;
    int32_t thermal_err;   /* set when dt_thermal / dt_mech <= 1 (Heat.c:121-126) */
    int32_t nonfinite;     /* set when a velocity or dt stops being finite */
    double topoGradMax;    /* running max of |grad y| on the top surface
                              (Context.c:759-764: never reset)            */
};

struct DevBC {
    uint32_t mask;
    double val[3];
};

/* n / d for 0 <= n < 2^31 by a multiply-high, an add and a shift
   (T. Granlund and P. Montgomery, "Division by invariant integers using
   multiplication", PLDI 1994, sec. 4): l = ceil(log2 d),
   m = floor(2^32 (2^l - d) / d) + 1 < 2^32, and n / d = (mulhi(m, n) + n) >> l,
   which cannot overflow while n < 2^31.  Replaces the ~15-instruction
   generic division for the mesh's index arithmetic. */
struct FastDiv {
    uint32_t m, l;
};
SNH_HD FastDiv fastdiv_make(uint32_t d) {
    uint32_t l = 0;
    while (((uint64_t)1 << l) < d) l++;
    FastDiv f;
    f.m = (uint32_t)(((((uint64_t)1 << l) - d) << 32) / d + 1);
    f.l = l;
    return f;
}
__device__ __forceinline__ uint32_t fastdiv(uint32_t n, FastDiv f) { return (__umulhi(n, f.m) + n) >> f.l; }

struct StepArgs {
    const double* __restrict__ xc;  /* coordinates of this step [3][N]     */
    const double* __restrict__ xp;  /* coordinates of the step before      */
    double* __restrict__ xn;        /* coordinates written by the node pass */
    /* stress, plastic strain and velocity are read from the first pointer
       and written through the second; the normal step passes the same
       buffer twice, a recorded replay passes consecutive tape slots */
    const double* v;                /* velocity [3][N]                     */
    double* v_out;
    const double* stress;           /* [6][10][E]                          */
    double* stress_out;
    const double* ps;               /* [10][E] or null                     */
    double* ps_out;
    double* __restrict__ visc;      /* [10][E] or null                     */
    const uint8_t* __restrict__ mat;      /* [E] or null (material 0)      */
    const double* __restrict__ refvisc;   /* [E] or null                   */
    const double* __restrict__ cohoff;    /* [E] or null                   */
    const uint8_t* __restrict__ bc;       /* [N] BC code, 0 = none         */
    /* replay tape of the node pass (null when not recording): force before
       damping [3][N], inertial mass [N], sign(v)*sign(f) per component packed
       2 bits each [N] */
    double* tape_f;
    double* tape_m;
    uint8_t* tape_code;
    /* prescribed damping branch of this step (packed as tape_code), or null:
       sign(v) is replaced by code * sign(f), which gives the same bits when
       the code is the one the step would compute */
    const uint8_t* damp_sched;
    /* dump diagnostics of this step, or null: element scalars [5][E]
       (strain-rate invariant, stress invariant, hydrostatic pressure,
       volume-averaged plastic strain, unused) and the damped nodal force
       [3][N] */
    double* diag;
    double* force_out;
    /* SnacTemperature, all null when the plugin is off */
    double* temp;      /* [N] nodal temperature                              */
    double* temp0;     /* [N] nodal temperature of the step before           */
    double* tpre;      /* [N] temperature of this step before the BCs        */
    double* hflux;     /* [3][10][E] heat flux                               */
    double* tstress;   /* [10][E] thermal stress, null unless computed       */
    double* dens;      /* [10][E] tet density of the last Stress call        */
    const uint8_t* tbc;   /* [N] temperature-condition code, 0 = none        */
    const double* tbcv;   /* value of each code                              */
    const DevMaterial* __restrict__ mats;  /* material table of this context */
    const DevBC* __restrict__ bcs;         /* velocity-condition table        */
    double* __restrict__ part_min;
    double* __restrict__ part_vmax;
    DevScalars* __restrict__ scal;
    int32_t I, J, K;        /* node counts   */
    int32_t EI, EJ, EK;     /* element counts */
    int32_t N, E;
    int32_t dtType;
    int32_t node_band;      /* k-planes per band of the node passes' order, 0 = plain */
    int32_t node_kc;        /* node planes per workgroup of the tiled node pass */
    double gravity, demf, strain_inert;
    double c_wave;          /* the wave dt type's speed of sound: max Vp over the elements */
    /* divisions by I, J, EI, EJ, node_band and J * node_band */
    FastDiv dI, dJ, dEI, dEJ, dKB, dPer;
};

/* kinematics of one tetrahedron from its 4 node coordinates and velocities,
   in TetraToNode order */
struct TetKin {
    double V;
    double sr[6];   /* raw strain rate 00,11,22,01,02,12 (StrainRate.c:65-184) */
    double w[3];    /* rotation 01,02,12 (UpdateElement.c:145-219)            */
};

/* cross product of face (p1,p2,p3) as Tetrahedra_SurfaceArea/Normal compute it */
SNH_HD void face_cross(const double* c1, const double* c2, const double* c3, double& a, double& b, double& c) {
    double x2 = c2[0] - c1[0], y2 = c2[1] - c1[1], z2 = c2[2] - c1[2];
    double x3 = c3[0] - c1[0], y3 = c3[1] - c1[1], z3 = c3[2] - c1[2];
    a = y2 * z3 - z2 * y3;
    b = z2 * x3 - x2 * z3;
    c = x2 * y3 - y2 * x3;
}

/* Workgroups are handed to the 8 XCDs round-robin (block b runs on XCD
   b % 8), each with its own L2.  Renumber so that every XCD works through one
   contiguous range of the mesh: elements sharing nodes then share an L2.
   (b, nb): a block of a range of nb blocks that starts on a multiple of 8
   (a member of a batched ensemble launch). */
__device__ __forceinline__ int xcd_block(const int b, const int nb) {
    const int x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + i;
}
__device__ __forceinline__ int xcd_block() { return xcd_block((int)blockIdx.x, (int)gridDim.x); }

/* Tetrahedra_Volume (Tetrahedra.c:44-67).  Its cross product is the one of
   face 0 (nodes 0,1,2): x3*y2 == y2*x3 exactly. */
SNH_HD double tet_volume_from_face0(const double* c1, const double* c4, double a0, double b0, double c0) {
    double x4 = c4[0] - c1[0], y4 = c4[1] - c1[1], z4 = c4[2] - c1[2];
    double v = snh_div6(x4 * a0 + y4 * b0 + z4 * c0);   /* == (...)/6.0 */
    return fabs(v);
}

SNH_HD double tet_volume(const double* c1, const double* c2, const double* c3, const double* c4) {
    double a, b, c;
    face_cross(c1, c2, c3, a, b, c);
    return tet_volume_from_face0(c1, c4, a, b, c);
}

/* geometry + velocity-gradient terms of one tet (UpdateElement.c:84-219,
   StrainRate.c:65-184): node k pairs with face (k+1)%4 */
SNH_HD void tet_kinematics(const double X[4][3], const double U[4][3], TetKin& o) {
    double n[4][3], A[4];
    double a0 = 0, b0 = 0, c0 = 0;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        double a, b, c;
        face_cross(X[kFaceToNode[f][0]], X[kFaceToNode[f][1]], X[kFaceToNode[f][2]], a, b, c);
        if (f == 0) {
            a0 = a;
            b0 = b;
            c0 = c;
        }
        double s = sqrt(a * a + b * b + c * c);
        const double ys = 1.0 / s;
        A[f] = s / 2.0;
        n[f][0] = snh_div_by(a, s, ys);   /* == a / s */
        n[f][1] = snh_div_by(b, s, ys);
        n[f][2] = snh_div_by(c, s, ys);
    }
    const double V = tet_volume_from_face0(X[0], X[3], a0, b0, c0);
    o.V = V;
#define T_(k, p, q) (U[k][p] * n[((k) + 1) & 3][q] * A[((k) + 1) & 3])
    /* products shared between strain rate and rotation */
    const double s00 = T_(0, 0, 0) + T_(1, 0, 0) + T_(2, 0, 0) + T_(3, 0, 0);
    const double s11 = T_(0, 1, 1) + T_(1, 1, 1) + T_(2, 1, 1) + T_(3, 1, 1);
    const double s22 = T_(0, 2, 2) + T_(1, 2, 2) + T_(2, 2, 2) + T_(3, 2, 2);
    const double s01 = T_(0, 0, 1) + T_(1, 0, 1) + T_(2, 0, 1) + T_(3, 0, 1);
    const double s02 = T_(0, 0, 2) + T_(1, 0, 2) + T_(2, 0, 2) + T_(3, 0, 2);
    const double s12 = T_(0, 1, 2) + T_(1, 1, 2) + T_(2, 1, 2) + T_(3, 1, 2);
    /* x / 3.0 / V as snh_div_by(snh_div3(x), V, RN(1/V)): the same two
       correctly rounded quotients */
    const double yV = 1.0 / V;
    o.sr[0] = snh_div_by(snh_div3(-1.0 * s00), V, yV);
    o.sr[1] = snh_div_by(snh_div3(-1.0 * s11), V, yV);
    o.sr[2] = snh_div_by(snh_div3(-1.0 * s22), V, yV);
    o.sr[3] = snh_div_by(snh_div3(-0.5 * (s01 + T_(0, 1, 0) + T_(1, 1, 0) + T_(2, 1, 0) + T_(3, 1, 0))), V, yV);
    o.sr[4] = snh_div_by(snh_div3(-0.5 * (s02 + T_(0, 2, 0) + T_(1, 2, 0) + T_(2, 2, 0) + T_(3, 2, 0))), V, yV);
    o.sr[5] = snh_div_by(snh_div3(-0.5 * (s12 + T_(0, 2, 1) + T_(1, 2, 1) + T_(2, 2, 1) + T_(3, 2, 1))), V, yV);
    const double hv = -(yV * 0.5);   /* RN(0.5/V) == RN(1/V)/2 */
    o.w[0] = snh_div3(hv * (s01 - T_(0, 1, 0) - T_(1, 1, 0) - T_(2, 1, 0) - T_(3, 1, 0)));
    o.w[1] = snh_div3(hv * (s02 - T_(0, 2, 0) - T_(1, 2, 0) - T_(2, 2, 0) - T_(3, 2, 0)));
    o.w[2] = snh_div3(hv * (s12 - T_(0, 2, 1) - T_(1, 2, 1) - T_(2, 2, 1) - T_(3, 2, 1)));
#undef T_
}

/* in-place, in-order Jaumann rotation of a tet stress (UpdateElement.c:245-265) */
SNH_HD void rotate_stress(double st[6], const double w[3], double dt) {
    const double w01 = w[0], w02 = w[1], w12 = w[2];
    st[0] += (2.0 * w01 * st[3] + 2.0 * w02 * st[4]) * dt;
    st[1] += (2.0 * (-1.0 * w01) * st[3] + 2.0 * w12 * st[5]) * dt;
    st[2] += (2.0 * (-1.0 * w02) * st[4] + 2.0 * (-1.0 * w12) * st[5]) * dt;
    st[3] += (w01 * (st[1] - st[0])) * dt + (w02 * st[5] + w12 * st[4]) * dt;
    st[4] += (w02 * (st[2] - st[0])) * dt + (w01 * st[5] - w12 * st[3]) * dt;
    st[5] += (w12 * (st[2] - st[1])) * dt + ((-1.0 * w01) * st[4] - w02 * st[3]) * dt;
}

}  // namespace snh

#endif
