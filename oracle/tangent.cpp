/*
 * tangent.cpp -- FP64 tangent-linear model of SNAC's explicit step on the CPU.
 *
 * TEST INFRASTRUCTURE ONLY: the reference the GPU adjoint (include/
 * snac_hip.h, snh_adjoint_*) is tested against.  Nothing in the product links
 * or calls it.
 *
 * It restates the step of the reference in the order of AbstractContext_Step
 * (Context.c:1063-1268): strain rate (StrainRate.c:65-305), strain and
 * constitutive update (Stress.c:53-172; elastic/Constitutive.c:40-66,
 * maxwell/Constitutive.c:47-184 with a linear viscosity law,
 * plastic_BI/Constitutive.c:59-238 with constant friction and dilation
 * angles), force and inertial mass (Force.c:87-183, dynamic dt), damping,
 * velocity, BC and coordinate update (UpdateNode.c:60-158), and the Jaumann
 * rotation of the stress with the new mesh (UpdateElement.c:145-265).  Every
 * quantity is a dual number, so one run gives the derivative of the final
 * coordinates along one direction of the parameter field.  The discrete
 * decisions are imposed from a record of the reference-faithful CPU oracle
 * (oracle/snac_oracle.c: orc_record_decisions): dt and speed of sound of each
 * step, sign(v) sign(f) of each damped node component and the plastic branch
 * of each tet.  Written independently of the GPU reverse kernels: geometry
 * from the face cross products a (n A = a / 2); the return from the
 * eigenpairs of the trial stress's values (the oracle's restatement of the
 * reference's solver, whose quirks the trajectory depends on) with its
 * derivative in eigenprojector form, written independently of the GPU
 * reverse kernel's eigenbasis form.  A second form of the return, the
 * closed-form eigenvalues (trigonometric solution of the characteristic
 * cubic) with the dual-number Sylvester assembly, checks the first where the
 * principal stresses are well apart.  Links liboracle.so.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../include/snac_hip.h"

#include "snac_oracle.h"

namespace {

struct Dual {
    double v, d;
    Dual() : v(0.0), d(0.0) {}
    Dual(double a) : v(a), d(0.0) {}
    Dual(double a, double b) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline double val(Dual a) { return a.v; }
inline Dual dsqrt(Dual a) {
    const double s = sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}
inline Dual dabs(Dual a) { return a.v < 0.0 ? -a : a; }
inline Dual dexp(Dual a) {
    const double e = exp(a.v);
    return Dual(e, a.d * e);
}
inline Dual dcos(Dual a) { return Dual(cos(a.v), -a.d * sin(a.v)); }
inline Dual dacos(Dual a) {
    double x = a.v;
    if (x >= 1.0) return Dual(0.0, 0.0);
    if (x <= -1.0) return Dual(M_PI, 0.0);
    return Dual(acos(x), -a.d / sqrt(1.0 - x * x));
}

/* the reference's tables (Snac/libSnac/src/TetrahedraTables.c:42-84) and
   hex-local node offsets (HexaEL.c:462-469) */
const int TetraToNode[10][4] = {{7, 2, 3, 0}, {0, 5, 4, 7}, {2, 5, 1, 0}, {5, 2, 6, 7}, {2, 5, 0, 7},
                                {3, 4, 7, 6}, {4, 3, 0, 1}, {6, 1, 2, 3}, {1, 6, 5, 4}, {1, 3, 6, 4}};
const int FaceToNode[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {3, 1, 0}};
const int NodeToTetra[8][5] = {{3, 5, 7, 8, 9}, {0, 1, 3, 4, 5}, {1, 2, 3, 4, 8}, {1, 5, 6, 8, 9},
                               {0, 2, 3, 4, 7}, {0, 5, 6, 7, 9}, {2, 6, 7, 8, 9}, {0, 1, 2, 4, 6}};
const int NodeToSurface[8][5] = {{3, 0, 1, 2, 3}, {1, 0, 0, 0, 3}, {2, 2, 1, 2, 3}, {3, 2, 1, 0, 0},
                                 {2, 1, 2, 1, 3}, {3, 1, 2, 0, 2}, {3, 0, 2, 1, 1}, {0, 1, 0, 3, 3}};
const int HexOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

void cross_face(const Dual* p1, const Dual* p2, const Dual* p3, Dual a[3]) {
    const Dual u0 = p2[0] - p1[0], u1 = p2[1] - p1[1], u2 = p2[2] - p1[2];
    const Dual w0 = p3[0] - p1[0], w1 = p3[1] - p1[1], w2 = p3[2] - p1[2];
    a[0] = u1 * w2 - u2 * w1;
    a[1] = u2 * w0 - u0 * w2;
    a[2] = u0 * w1 - u1 * w0;
}

/* volume |(x3 - x0) . a0| / 6 and the 4 face cross products of a tet */
Dual tet_geometry(const Dual X[4][3], Dual a[4][3]) {
    for (int f = 0; f < 4; f++) cross_face(X[FaceToNode[f][0]], X[FaceToNode[f][1]], X[FaceToNode[f][2]], a[f]);
    const Dual det = (X[3][0] - X[0][0]) * a[0][0] + (X[3][1] - X[0][1]) * a[0][1] + (X[3][2] - X[0][2]) * a[0][2];
    return dabs(det) / 6.0;
}

/* velocity gradient L = -(1/(6V)) sum_k u_k (x) a_(k+1)%4 : strain rate
   sym(L) (00,11,22,01,02,12) and spin skew(L) (01,02,12) */
void tet_rates(const Dual U[4][3], const Dual a[4][3], Dual V, Dual sr[6], Dual w[3]) {
    Dual L[3][3];
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
            Dual g = 0.0;
            for (int k = 0; k < 4; k++) g += U[k][p] * a[(k + 1) & 3][q];
            L[p][q] = -g / (6.0 * V);
        }
    sr[0] = L[0][0];
    sr[1] = L[1][1];
    sr[2] = L[2][2];
    sr[3] = 0.5 * (L[0][1] + L[1][0]);
    sr[4] = 0.5 * (L[0][2] + L[2][0]);
    sr[5] = 0.5 * (L[1][2] + L[2][1]);
    w[0] = 0.5 * (L[0][1] - L[1][0]);
    w[1] = 0.5 * (L[0][2] - L[2][0]);
    w[2] = 0.5 * (L[1][2] - L[2][1]);
}

/* in-order Jaumann update (UpdateElement.c:245-265) */
void rotate(Dual s[6], const Dual w[3], double dt) {
    const Dual w01 = w[0], w02 = w[1], w12 = w[2];
    s[0] += (2.0 * w01 * s[3] + 2.0 * w02 * s[4]) * dt;
    s[1] += (-2.0 * w01 * s[3] + 2.0 * w12 * s[5]) * dt;
    s[2] += (-2.0 * w02 * s[4] - 2.0 * w12 * s[5]) * dt;
    s[3] += w01 * (s[1] - s[0]) * dt + (w02 * s[5] + w12 * s[4]) * dt;
    s[4] += w02 * (s[2] - s[0]) * dt + (w01 * s[5] - w12 * s[3]) * dt;
    s[5] += w12 * (s[2] - s[1]) * dt + (-w01 * s[4] - w02 * s[3]) * dt;
}

/* eigenvalues (ascending) of a symmetric 3x3 in closed form */
void eig3(const Dual S[3][3], Dual lam[3]) {
    const Dual p1 = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
    const Dual q = (S[0][0] + S[1][1] + S[2][2]) / 3.0;
    const Dual b0 = S[0][0] - q, b1 = S[1][1] - q, b2 = S[2][2] - q;
    const Dual p2 = b0 * b0 + b1 * b1 + b2 * b2 + 2.0 * p1;
    const Dual p = dsqrt(p2 / 6.0);
    Dual B[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B[i][j] = (S[i][j] - (i == j ? q : Dual(0.0))) / p;
    const Dual det = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) - B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                     B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
    const Dual phi = dacos(det / 2.0) / 3.0;
    lam[2] = q + 2.0 * p * dcos(phi);
    lam[0] = q + 2.0 * p * dcos(phi + 2.0 * M_PI / 3.0);
    lam[1] = 3.0 * q - lam[0] - lam[2];
}

/* Eigenpairs of the trial stress's values, ascending, with the projectors
   P_k = v_k v_k^T, from the CPU oracle's restatement of the reference's
   solver (Householder tridiagonalisation and QL iterations with the
   reference's int abs(), plastic_BI/Constitutive.c:240-510).  Its int abs()
   takes off-diagonals below 1 Pa as converged; where two principal stresses
   lie within about 1 Pa of each other, the eigenvectors it returns, and with
   them the return's result, depend on that, so the tangent model takes them
   from the same solver to follow the reference's trajectory. */
void reference_eigenpairs(const double A[3][3], double w[3], double Pk[3][3][3]) {
    double V[3][3];
    orc_principal_stresses(A, w, V);
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Pk[k][i][j] = V[i][k] * V[j][k];
}

/* Eigenpairs of a symmetric 3x3 (values) by cyclic Jacobi rotations,
   ascending, with the projectors P_k = v_k v_k^T.  Rotations drive every
   off-diagonal to zero (or below 1e-18 of its diagonal pair), so the
   eigenpairs are smooth functions of the matrix wherever its eigenvalues are
   distinct: the value map whose finite differences check the dual-number
   derivative. */
void jacobi3(const double A0[3][3], double w[3], double Pk[3][3][3]) {
    double a[3][3], V[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            a[i][j] = A0[i][j];
            V[i][j] = i == j ? 1.0 : 0.0;
        }
    static const int PQ[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 64; sweep++) {
        bool rotated = false;
        for (const auto& pq : PQ) {
            const int p = pq[0], q = pq[1];
            if (a[p][q] == 0.0) continue;
            if (fabs(a[p][q]) <= 1e-18 * (fabs(a[p][p]) + fabs(a[q][q]))) {
                a[p][q] = a[q][p] = 0.0;
                continue;
            }
            rotated = true;
            /* t = tan of the angle that zeroes a_pq, the smaller root of
               t^2 + 2 th t - 1 = 0, th = (a_qq - a_pp) / (2 a_pq) */
            const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
            const double t = fabs(th) > 1e150 ? 0.5 / th
                                              : (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
            for (int k = 0; k < 3; k++) {   /* A J */
                const double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - sn * akq;
                a[k][q] = sn * akp + c * akq;
            }
            for (int k = 0; k < 3; k++) {   /* J^T (A J) */
                const double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - sn * aqk;
                a[q][k] = sn * apk + c * aqk;
            }
            a[p][q] = a[q][p] = 0.0;
            for (int k = 0; k < 3; k++) {   /* V J */
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - sn * vkq;
                V[k][q] = sn * vkp + c * vkq;
            }
        }
        if (!rotated) break;
    }
    int order[3] = {0, 1, 2};
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++)
            if (a[order[j]][order[j]] < a[order[i]][order[i]]) {
                const int t = order[i];
                order[i] = order[j];
                order[j] = t;
            }
    for (int k = 0; k < 3; k++) {
        const int o = order[k];
        w[k] = a[o][o];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Pk[k][i][j] = V[i][o] * V[j][o];
    }
}

/* The returned stress sum_k g_k P_k and its derivative in projector form:
     d(sum_k g_k P_k) = sum_k dg_k P_k
                        + sum_{k<l} G_kl (P_k dS P_l + P_l dS P_k),
     G_kl = (g_k - g_l) / (lam_k - lam_l),
   the second sum being sum_k g_k dP_k (dP_k = sum_{l != k} (P_k dS P_l +
   P_l dS P_k) / (lam_k - lam_l)).  G_kl is the factor by which the return
   scales the rotation of the eigenbasis; with clip it is held to [0, 1] (the
   corner clamp of the adjoint, SNH_ADJ_CORNER_CLAMP), and it is 1 where the
   two eigenvalues coincide to 1e-14 of their size. */
void projector_return(const Dual S[3][3], const Dual lam[3], const Dual g[3], const double Pk[3][3][3], bool clip,
                      Dual out[3][3]) {
    double dS[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) dS[i][j] = S[i][j].d;
    double v[3][3], d[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            v[i][j] = g[0].v * Pk[0][i][j] + g[1].v * Pk[1][i][j] + g[2].v * Pk[2][i][j];
            d[i][j] = g[0].d * Pk[0][i][j] + g[1].d * Pk[1][i][j] + g[2].d * Pk[2][i][j];
        }
    for (int k = 0; k < 3; k++)
        for (int l = k + 1; l < 3; l++) {
            const double gap = lam[k].v - lam[l].v;
            double G = 1.0;
            if (fabs(gap) > 1e-14 * (fabs(lam[k].v) + fabs(lam[l].v))) G = (g[k].v - g[l].v) / gap;
            if (clip) G = G < 0.0 ? 0.0 : (G > 1.0 ? 1.0 : G);
            double A[3][3], B[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    A[i][j] = 0.0;
                    B[i][j] = 0.0;
                    for (int r = 0; r < 3; r++) {
                        A[i][j] += Pk[k][i][r] * dS[r][j];
                        B[i][j] += Pk[l][i][r] * dS[r][j];
                    }
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double t = 0.0;
                    for (int r = 0; r < 3; r++) t += A[i][r] * Pk[l][r][j] + B[i][r] * Pk[k][r][j];
                    d[i][j] += G * t;
                }
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i][j] = Dual(v[i][j], d[i][j]);
}

struct Mat {
    uint32_t rheology;
    double lambda, mu, rho, lam2mu;
    uint32_t nseg;
    double plstrain[SNH_MAX_SEGMENTS + 1], coh[SNH_MAX_SEGMENTS + 1];
    double phi, psi, ten_off;
    double vis_min, vis_max, refvisc;
};

}  // namespace

extern "C" {

typedef struct tg_problem {
    int32_t I, J, K;              /* node counts */
    const double* x0;             /* [N][3] coordinates after TimeStepZero */
    const double* v0;             /* [N][3] */
    const double* stress0;        /* [E][10][6] */
    const double* ps0;            /* [E][10] or NULL */
    const uint8_t* mat;           /* [E] or NULL */
    const snh_material* materials;
    int32_t nmat;
    const double* refvisc;        /* [E] or NULL */
    const double* cohoff;         /* [E] or NULL */
    const uint8_t* bc_mask;       /* [N] */
    const double* bc_value;       /* [N][3] */
    double gravity, demf;
    int32_t param;                /* SNH_PARAM_* */
    int32_t flags;                /* SNH_ADJ_CORNER_CLAMP, TG_SYLVESTER, TG_JACOBI */
    void* follow;                 /* orc_context* at the start state, or NULL */
} tg_problem;

/* flags bit of tg_problem beside SNH_ADJ_CORNER_CLAMP: the Mohr-Coulomb
   return through the closed-form eigenvalues (trigonometric solution of the
   characteristic cubic) and the dual-number Sylvester assembly, in place of
   the reference's eigenpairs and the projector form; a cross-check of the
   two where the principal stresses are well apart.  Not combined with the
   clamp. */
#define TG_SYLVESTER 0x100
/* flags bit: eigenpairs by Jacobi rotations (a smooth value map) in place of
   the reference's solver */
#define TG_JACOBI 0x200

typedef struct tg_decisions {
    int32_t nsteps;
    const double* dt;             /* [nsteps] */
    const double* c;              /* [nsteps] speed of sound */
    const int8_t* damp;           /* [nsteps][N][3] */
    const int8_t* branch;         /* [nsteps][E][10] */
} tg_decisions;

/* Run nsteps from the given state with the parameter field moved along dir
   ([E], NULL = no motion).  x_out: final coordinates [N][3]; dx_out: their
   derivative along dir [N][3].  Returns 0, or -1 on a bad argument. */
int tg_run(const tg_problem* P, const tg_decisions* D, const double* dir, double* x_out, double* dx_out) {
    if (!P || !D || !x_out || !dx_out || P->nmat < 1) return -1;
    if ((P->flags & TG_SYLVESTER) && (P->flags & SNH_ADJ_CORNER_CLAMP)) return -1;
    const int I = P->I, J = P->J, K = P->K;
    const int EI = I - 1, EJ = J - 1, EK = K - 1;
    const size_t N = (size_t)I * J * K, E = (size_t)EI * EJ * EK;
    std::vector<Mat> mats(P->nmat);
    for (int i = 0; i < P->nmat; i++) {
        const snh_material& s = P->materials[i];
        Mat& m = mats[i];
        m.rheology = s.rheology;
        m.lambda = s.lambda;
        m.mu = s.mu;
        m.rho = s.phsDensity * (1.0 - s.alpha * (0.0 - s.reftemp));
        m.lam2mu = s.lambda + 2.0 * s.mu;
        m.nseg = s.nsegments;
        for (uint32_t k = 0; k <= s.nsegments && k <= SNH_MAX_SEGMENTS; k++) {
            m.plstrain[k] = s.plstrain[k];
            m.coh[k] = s.cohesion[k];
        }
        m.phi = s.frictionAngle[0];
        m.psi = s.dilationAngle[0];
        m.ten_off = s.ten_off;
        m.vis_min = s.vis_min;
        m.vis_max = s.vis_max;
        m.refvisc = s.refvisc;
    }
    std::vector<Dual> x(3 * N), xp(3 * N), v(3 * N), st(60 * E), ps(10 * E), rv(E), co(E);
    for (size_t i = 0; i < 3 * N; i++) {
        x[i] = P->x0[i];
        xp[i] = P->x0[i];
        v[i] = P->v0[i];
    }
    for (size_t i = 0; i < 60 * E; i++) st[i] = P->stress0[i];
    for (size_t i = 0; i < 10 * E; i++) ps[i] = P->ps0 ? P->ps0[i] : 0.0;
    for (size_t e = 0; e < E; e++) {
        const Mat& m = mats[P->mat ? P->mat[e] : 0];
        const double d = dir ? dir[e] : 0.0;
        if (P->param == SNH_PARAM_LOG_VISCOSITY) {
            const double r = P->refvisc ? P->refvisc[e] : m.refvisc;
            rv[e] = dexp(Dual(log(r), d));
            co[e] = P->cohoff ? P->cohoff[e] : 0.0;
        } else {
            rv[e] = P->refvisc ? P->refvisc[e] : m.refvisc;
            co[e] = Dual(P->cohoff ? P->cohoff[e] : 0.0, d);
        }
    }
    auto nid = [&](int i, int j, int k) { return (size_t)i + (size_t)I * ((size_t)j + (size_t)J * k); };
    auto hex = [&](size_t e, int q) {
        const int ei = (int)(e % EI), ej = (int)((e / EI) % EJ), ek = (int)(e / ((size_t)EI * EJ));
        return nid(ei + HexOff[q][0], ej + HexOff[q][1], ek + HexOff[q][2]);
    };
    std::vector<Dual> f(3 * N), mass(N), xn(3 * N);
    std::vector<double> fx, fv, fs, fp;
    orc_context* follow = (orc_context*)P->follow;
    if (follow) {
        if (orc_node_count(follow) != N || orc_element_count(follow) != E) return -1;
        fx.resize(3 * N);
        fv.resize(3 * N);
        fs.resize(60 * E);
        fp.resize(10 * E);
    }
    for (int step = 0; step < D->nsteps; step++) {
        const double dt = D->dt[step], cs = D->c[step];
        if (follow) {
            /* linearise about the oracle's state of this step: the values
               are the oracle's, the derivatives carry on */
            snh_state os;
            memset(&os, 0, sizeof os);
            os.coord = fx.data();
            os.velocity = fv.data();
            os.stress = fs.data();
            os.plastic_strain = fp.data();
            if (orc_get_state(follow, &os) != 0) return -1;
            for (size_t i = 0; i < 3 * N; i++) {
                x[i].v = fx[i];
                v[i].v = fv[i];
            }
            for (size_t i = 0; i < 60 * E; i++) st[i].v = fs[i];
            for (size_t i = 0; i < 10 * E; i++) ps[i].v = fp[i];
        }
        const int8_t* br = D->branch ? D->branch + (size_t)step * 10 * E : nullptr;
        const int8_t* dm = D->damp + (size_t)step * 3 * N;
        /* ---- strain rate, strain, constitutive, trace averaging ---- */
        for (size_t e = 0; e < E; e++) {
            const Mat& m = mats[P->mat ? P->mat[e] : 0];
            Dual Vt[10], sr[10][6], w[10][3], tr[10];
            Dual EV = 0.0, partial = 0.0;
            for (int t = 0; t < 10; t++) {
                Dual X[4][3], U[4][3], a[4][3];
                for (int q = 0; q < 4; q++) {
                    const size_t n = hex(e, TetraToNode[t][q]);
                    for (int d = 0; d < 3; d++) {
                        X[q][d] = x[3 * n + d];
                        U[q][d] = v[3 * n + d];
                    }
                }
                Vt[t] = tet_geometry(X, a);
                tet_rates(U, a, Vt[t], sr[t], w[t]);
                tr[t] = (sr[t][0] + sr[t][1] + sr[t][2]) / 3.0;
                EV += Vt[t];
                partial += tr[t] * Vt[t];
            }
            const Dual srcorr = partial / EV;
            /* softened cohesion carried along the tets (plastic_BI :78-132) */
            Dual pcoh = 30.0e6;
            Dual pst = 0.0, trs[10];
            for (int t = 0; t < 10; t++) {
                Dual* s = &st[(e * 10 + t) * 6];
                Dual sn[6];
                for (int c = 0; c < 6; c++) sn[c] = sr[t][c] * dt;
                for (int c = 0; c < 3; c++) sn[c] = (sr[t][c] + (srcorr - tr[t])) * dt;
                const Dual trs_n = sn[0] + sn[1] + sn[2];
                if (m.rheology & SNH_RHEOLOGY_PLASTIC) {
                    Dual& pst_t = ps[e * 10 + t];
                    for (uint32_t i = 0; i < m.nseg; i++)
                        if (pst_t.v >= m.plstrain[i] && pst_t.v <= m.plstrain[i + 1])
                            pcoh = m.coh[i] + (m.coh[i + 1] - m.coh[i]) / (m.plstrain[i + 1] - m.plstrain[i]) *
                                                  (pst_t - m.plstrain[i]);
                    const Dual coh = pcoh + co[e];
                    const double degrad = M_PI / 180.0;
                    Dual tc = m.ten_off;
                    if (m.phi > 0.0 && m.ten_off == 0.0) tc = coh / tan(m.phi * degrad);
                    for (int c = 0; c < 6; c++) s[c] += 2.0 * m.mu * sn[c];
                    for (int c = 0; c < 3; c++) s[c] += m.lambda * trs_n;
                    const int b = br ? br[e * 10 + t] : 0;
                    if (b) {
                        Dual S[3][3] = {{s[0], s[3], s[4]}, {s[3], s[1], s[5]}, {s[4], s[5], s[2]}};
                        Dual lam[3];
                        double Pk[3][3][3];
                        const bool sylvester = (P->flags & TG_SYLVESTER) != 0;
                        if (sylvester) {
                            eig3(S, lam);
                        } else {
                            double Sv[3][3], w[3];
                            for (int i = 0; i < 3; i++)
                                for (int j = 0; j < 3; j++) Sv[i][j] = S[i][j].v;
                            if (P->flags & TG_JACOBI) jacobi3(Sv, w, Pk);
                            else reference_eigenpairs(Sv, w, Pk);
                            /* d lam_k = P_k : dS */
                            for (int k = 0; k < 3; k++) {
                                double dl = 0.0;
                                for (int i = 0; i < 3; i++)
                                    for (int j = 0; j < 3; j++) dl += Pk[k][i][j] * S[i][j].d;
                                lam[k] = Dual(w[k], dl);
                            }
                        }
                        const double sphi = sin(m.phi * degrad), spsi = sin(m.psi * degrad);
                        const double anphi = (1.0 + sphi) / (1.0 - sphi), anpsi = (1.0 + spsi) / (1.0 - spsi);
                        const double a1 = m.lambda + 2.0 * m.mu, a2 = m.lambda;
                        Dual alam, g[3];
                        double dep[3];
                        if (b == 1) {
                            const Dual fs = lam[0] - lam[2] * anphi + 2.0 * coh * sqrt(anphi);
                            alam = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi);
                            g[0] = lam[0] - alam * (a1 - a2 * anpsi);
                            g[1] = lam[1] - alam * a2 * (1.0 - anpsi);
                            g[2] = lam[2] - alam * (a2 - a1 * anpsi);
                            dep[0] = 1.0;
                            dep[1] = 0.0;
                            dep[2] = -anpsi;
                        } else {
                            alam = (lam[2] - tc) / a1;
                            g[0] = lam[0] - alam * a2;
                            g[1] = lam[1] - alam * a2;
                            g[2] = lam[2] - alam * a1;
                            dep[0] = 0.0;
                            dep[1] = 0.0;
                            dep[2] = 1.0;
                        }
                        const double dm3 = (dep[0] + dep[1] + dep[2]) / 3.0;
                        const double Kf = sqrt(0.5 * ((dep[0] - dm3) * (dep[0] - dm3) + (dep[1] - dm3) * (dep[1] - dm3) +
                                                      (dep[2] - dm3) * (dep[2] - dm3) + dm3 * dm3));
                        pst_t = pst_t + Kf * dabs(alam);
                        Dual out[3][3];
                        if (sylvester) {
                            /* sum_k g_k P_k, P_k = prod_{l != k} (S - lam_l) / (lam_k - lam_l) */
                            for (int i = 0; i < 3; i++)
                                for (int j = 0; j < 3; j++) out[i][j] = 0.0;
                            for (int k = 0; k < 3; k++) {
                                const int l1 = (k + 1) % 3, l2 = (k + 2) % 3;
                                Dual A1[3][3], A2[3][3];
                                for (int i = 0; i < 3; i++)
                                    for (int j = 0; j < 3; j++) {
                                        A1[i][j] = S[i][j] - (i == j ? lam[l1] : Dual(0.0));
                                        A2[i][j] = S[i][j] - (i == j ? lam[l2] : Dual(0.0));
                                    }
                                const Dual den = (lam[k] - lam[l1]) * (lam[k] - lam[l2]);
                                for (int i = 0; i < 3; i++)
                                    for (int j = 0; j < 3; j++) {
                                        Dual pr = 0.0;
                                        for (int r = 0; r < 3; r++) pr += A1[i][r] * A2[r][j];
                                        out[i][j] += g[k] * pr / den;
                                    }
                            }
                        } else {
                            projector_return(S, lam, g, Pk, (P->flags & SNH_ADJ_CORNER_CLAMP) != 0, out);
                        }
                        s[0] = out[0][0];
                        s[1] = out[1][1];
                        s[2] = out[2][2];
                        s[3] = out[0][1];
                        s[4] = out[0][2];
                        s[5] = out[1][2];
                    }
                } else if (m.rheology & SNH_RHEOLOGY_MAXWELL) {
                    Dual Xp[4][3], ap[4][3];
                    for (int q = 0; q < 4; q++) {
                        const size_t n = hex(e, TetraToNode[t][q]);
                        for (int d = 0; d < 3; d++) Xp[q][d] = xp[3 * n + d];
                    }
                    const Dual Vold = tet_geometry(Xp, ap);
                    const Dual te = Vt[t] / Vold - 1.0;
                    Dual visc = rv[e];
                    if (visc.v < m.vis_min) visc = m.vis_min;
                    if (visc.v > m.vis_max) visc = m.vis_max;
                    const Dual temp = m.mu / (2.0 * visc) * dt;
                    const Dual vic1 = 1.0 - temp, vic2 = 1.0 / (1.0 + temp);
                    const Dual ts3 = (s[0] + s[1] + s[2]) / 3.0, te3 = te / 3.0;
                    const double bulkm = m.lambda + 2.0 * m.mu / 3.0;
                    Dual d[3];
                    for (int i = 0; i < 3; i++) d[i] = ((s[i] - ts3) * vic1 + 2.0 * m.mu * (sn[i] - te3)) * vic2;
                    for (int c = 3; c < 6; c++) s[c] = (s[c] * vic1 + 2.0 * m.mu * sn[c]) * vic2;
                    for (int i = 0; i < 3; i++) s[i] = d[i] + ts3 + bulkm * te;
                } else {
                    for (int c = 0; c < 6; c++) s[c] += 2.0 * m.mu * sn[c];
                    for (int c = 0; c < 3; c++) s[c] += m.lambda * trs_n;
                }
                trs[t] = (s[0] + s[1] + s[2]) / 3.0;
                pst += trs[t] * Vt[t];
            }
            const Dual avg = pst / EV;
            for (int t = 0; t < 10; t++) {
                Dual* s = &st[(e * 10 + t) * 6];
                for (int c = 0; c < 3; c++) s[c] += avg - trs[t];
            }
        }
        /* ---- force and inertial mass ---- */
        for (size_t n = 0; n < N; n++) {
            f[3 * n] = f[3 * n + 1] = f[3 * n + 2] = 0.0;
            mass[n] = 0.0;
        }
        for (size_t n = 0; n < N; n++) {
            const int i = (int)(n % I), j = (int)((n / I) % J), k = (int)(n / ((size_t)I * J));
            for (int slot = 0; slot < 8; slot++) {
                const int ei = i - 1 + (slot & 1), ej = j - 1 + ((slot >> 1) & 1), ek = k - 1 + ((slot >> 2) & 1);
                if (ei < 0 || ej < 0 || ek < 0 || ei >= EI || ej >= EJ || ek >= EK) continue;
                const size_t e = (size_t)ei + (size_t)EI * ((size_t)ej + (size_t)EJ * ek);
                const Mat& m = mats[P->mat ? P->mat[e] : 0];
                for (int idx = 0; idx < 5; idx++) {
                    const int t = NodeToTetra[slot][idx], sf = NodeToSurface[slot][idx];
                    Dual X[4][3], a[4][3];
                    for (int q = 0; q < 4; q++) {
                        const size_t nn = hex(e, TetraToNode[t][q]);
                        for (int d = 0; d < 3; d++) X[q][d] = x[3 * nn + d];
                    }
                    const Dual V = tet_geometry(X, a);
                    const Dual* s = &st[(e * 10 + t) * 6];
                    f[3 * n + 0] += (s[0] * a[sf][0] + s[3] * a[sf][1] + s[4] * a[sf][2]) / 12.0;
                    f[3 * n + 1] += (s[3] * a[sf][0] + s[1] * a[sf][1] + s[5] * a[sf][2]) / 12.0;
                    f[3 * n + 2] += (s[4] * a[sf][0] + s[5] * a[sf][1] + s[2] * a[sf][2]) / 12.0;
                    f[3 * n + 1] -= P->gravity * m.rho * V / 8.0;
                    mass[n] += 0.25 * m.lam2mu / (cs * cs) * V;
                }
            }
        }
        /* ---- damping, velocity, BC, coordinates ---- */
        for (size_t n = 0; n < N; n++)
            for (int d = 0; d < 3; d++) {
                const double kd = 1.0 - P->demf * (double)dm[3 * n + d];
                Dual vv = v[3 * n + d] + dt * kd * f[3 * n + d] / mass[n];
                if (P->bc_mask[n] & (1u << d)) vv = P->bc_value[3 * n + d];
                v[3 * n + d] = vv;
                xn[3 * n + d] = x[3 * n + d] + dt * vv;
            }
        xp = x;
        x = xn;
        /* ---- Jaumann rotation with the new mesh (UpdateElement.c) ---- */
        for (size_t e = 0; e < E; e++)
            for (int t = 0; t < 10; t++) {
                Dual X[4][3], U[4][3], a[4][3], sr[6], w[3];
                for (int q = 0; q < 4; q++) {
                    const size_t n = hex(e, TetraToNode[t][q]);
                    for (int d = 0; d < 3; d++) {
                        X[q][d] = x[3 * n + d];
                        U[q][d] = v[3 * n + d];
                    }
                }
                const Dual V = tet_geometry(X, a);
                tet_rates(U, a, V, sr, w);
                rotate(&st[(e * 10 + t) * 6], w, dt);
            }
        if (follow && orc_step(follow, 1) != 0) return -1;
    }
    for (size_t i = 0; i < 3 * N; i++) {
        x_out[i] = x[i].v;
        dx_out[i] = x[i].d;
    }
    return 0;
}

}  /* extern "C" */
