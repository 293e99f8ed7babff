/*
 * snac_hip.hip -- gfx950 kernels and C ABI (include/snac_hip.h) of SNAC's
 * explicit time step, FP64 path.
 *
 * One step is four launches on the context's stream:
 *   k_elem   per element: geometry of the current mesh, the pending Jaumann
 *            rotation of the previous step (UpdateElement.c:145-265), strain
 *            rate (StrainRate.c), strain, constitutive update and trace
 *            averaging (Stress.c).  Geometry and velocity gradient are computed
 *            once and serve both the rotation and the strain rate, whose inputs
 *            are identical in the reference.
 *   k_node   per node: atomic-free gather over the 8 surrounding elements x 5
 *            tets in the reference's order (Force.c:87-183), damping, velocity,
 *            velocity BC and coordinate update (UpdateNode.c:60-158), and the
 *            block maximum of |v| over constrained nodes.
 *   k_geo    per element: tet volumes and face areas of the new mesh and the
 *            block minimum of V/A (UpdateElement.c:84-139).
 *   k_fin    one block: global min/max and the dynamic dt (Context.c:1208-1268).
 * dt stays on the device, so a run of steps needs no host synchronisation.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/snac_hip.h"
#include "snh_kernels.h"

using namespace snh;


#define ELEM_BLOCK 128
#define NODE_BLOCK 128
#define TOPO_MAX_I 4096   /* nodes along x the topography sweep holds in LDS */
#define GEO_BLOCK 256
#define FIN_BLOCK 1024

/* --------------------------------------------------------------------------
 * helpers
 * ------------------------------------------------------------------------ */
/* material of element e: with one material (no per-element ids) the table
   entry is uniform across the wave and its fields are scalar loads */
template <bool MM>
__device__ __forceinline__ const DevMaterial& mat_of(const StepArgs& a, int e) {
    if constexpr (MM)
        return a.mats[a.mat[e]];
    else
        return a.mats[0];
}

__device__ __forceinline__ int node_index(const StepArgs& a, int i, int j, int k) {
    return i + a.I * (j + a.J * k);
}

/* (ei, ej, ek) of element e and (i, j, k) of node n */
__device__ __forceinline__ void elem_ijk(const StepArgs& a, int e, int& ei, int& ej, int& ek) {
    const int q = (int)fastdiv((uint32_t)e, a.dEI);
    ek = (int)fastdiv((uint32_t)q, a.dEJ);
    ei = e - q * a.EI;
    ej = q - ek * a.EJ;
}
__device__ __forceinline__ void node_ijk(const StepArgs& a, int n, int& i, int& j, int& k) {
    const int q = (int)fastdiv((uint32_t)n, a.dI);
    k = (int)fastdiv((uint32_t)q, a.dJ);
    i = n - q * a.I;
    j = q - k * a.J;
}

/* load the 8 hex nodes of element (i,j,k): coords and velocities */
__device__ __forceinline__ void load_hex(const StepArgs& a, const double* __restrict__ x, int i, int j, int k,
                                         double X[8][3]) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
        int n = node_index(a, i + kHexOff[q][0], j + kHexOff[q][1], k + kHexOff[q][2]);
        X[q][0] = x[n];
        X[q][1] = x[a.N + n];
        X[q][2] = x[2 * a.N + n];
    }
}

__device__ __forceinline__ double block_min(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int nw = (blockDim.x + 63) >> 6;
    if (wid == 0) {
        v = lane < nw ? sh[lane] : INFINITY;
        for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    }
    return v;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int nw = (blockDim.x + 63) >> 6;
    if (wid == 0) {
        v = lane < nw ? sh[lane] : -INFINITY;
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    }
    return v;
}

/* --------------------------------------------------------------------------
 * eigen-decomposition of plastic_BI (Constitutive.c:240-510): Householder
 * tridiagonalisation and implicit QL for n = 3 with the reference's int abs()
 * (EISPACK tred2/tql2 as restated there).
 * ------------------------------------------------------------------------ */
__device__ void tred2_3(double d[3], double e[3], double V[3][3]) {
    const int n = 3;
#pragma unroll
    for (int j = 0; j < n; j++) d[j] = V[n - 1][j];
#pragma unroll
    for (int i = n - 1; i > 0; i--) {
        double scale = 0.0, h = 0.0;
#pragma unroll
        for (int k = 0; k < i; k++) scale = scale + snh_ref_iabs(d[k]);
        if (scale == 0.0) {
            e[i] = d[i - 1];
#pragma unroll
            for (int j = 0; j < i; j++) {
                d[j] = V[i - 1][j];
                V[i][j] = 0.0;
                V[j][i] = 0.0;
            }
        } else {
            double f, g, hh;
#pragma unroll
            for (int k = 0; k < i; k++) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            f = d[i - 1];
            g = sqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h = h - f * g;
            d[i - 1] = f - g;
#pragma unroll
            for (int j = 0; j < i; j++) e[j] = 0.0;
#pragma unroll
            for (int j = 0; j < i; j++) {
                f = d[j];
                V[j][i] = f;
                g = e[j] + V[j][j] * f;
#pragma unroll
                for (int k = j + 1; k <= i - 1; k++) {
                    g += V[k][j] * d[k];
                    e[k] += V[k][j] * f;
                }
                e[j] = g;
            }
            f = 0.0;
#pragma unroll
            for (int j = 0; j < i; j++) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            hh = f / (h + h);
#pragma unroll
            for (int j = 0; j < i; j++) e[j] -= hh * d[j];
#pragma unroll
            for (int j = 0; j < i; j++) {
                f = d[j];
                g = e[j];
#pragma unroll
                for (int k = j; k <= i - 1; k++) V[k][j] -= (f * e[k] + g * d[k]);
                d[j] = V[i - 1][j];
                V[i][j] = 0.0;
            }
        }
        d[i] = h;
    }
#pragma unroll
    for (int i = 0; i < n - 1; i++) {
        double h, g;
        V[n - 1][i] = V[i][i];
        V[i][i] = 1.0;
        h = d[i + 1];
        if (h != 0.0) {
#pragma unroll
            for (int k = 0; k <= i; k++) d[k] = V[k][i + 1] / h;
#pragma unroll
            for (int j = 0; j <= i; j++) {
                g = 0.0;
#pragma unroll
                for (int k = 0; k <= i; k++) g += V[k][i + 1] * V[k][j];
#pragma unroll
                for (int k = 0; k <= i; k++) V[k][j] -= g * d[k];
            }
        }
#pragma unroll
        for (int k = 0; k <= i; k++) V[k][i + 1] = 0.0;
    }
#pragma unroll
    for (int j = 0; j < n; j++) {
        d[j] = V[n - 1][j];
        V[n - 1][j] = 0.0;
    }
    V[n - 1][n - 1] = 1.0;
    e[0] = 0.0;
}

/* QL sweeps on a 3x3 tridiagonal; the column index m and the sweep loop are
   data dependent, so V is updated through a column-pair helper that keeps
   every array index static. */
__device__ __forceinline__ void rot_cols(double V[3][3], int i, double s, double c) {
    /* columns i and i+1, i in {0,1} */
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double vi = (i == 0) ? V[k][0] : V[k][1];
        double vi1 = (i == 0) ? V[k][1] : V[k][2];
        double h = vi1;
        double n1 = s * vi + c * h;
        double n0 = c * vi - s * h;
        if (i == 0) {
            V[k][1] = n1;
            V[k][0] = n0;
        } else {
            V[k][2] = n1;
            V[k][1] = n0;
        }
    }
}

__device__ __forceinline__ double sel3(const double a[3], int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }
__device__ __forceinline__ void set3(double a[3], int i, double v) {
    if (i == 0) a[0] = v;
    else if (i == 1) a[1] = v;
    else a[2] = v;
}

__device__ void tql2_3(double d[3], double e[3], double V[3][3]) {
    const int n = 3;
    double f = 0.0, tst1 = 0.0;
    const double eps = 2.220446049250313080847e-16; /* pow(2.0,-52.0) */
    e[0] = e[1];
    e[1] = e[2];
    e[2] = 0.0;
#pragma unroll
    for (int l = 0; l < n; l++) {
        int m = l;
        {
            double b = (double)snh_iadd_wrap(snh_ref_iabs(d[l]), snh_ref_iabs(e[l]));
            tst1 = (tst1 > b) ? tst1 : b;
        }
        while (m < n) {
            if (snh_ref_iabs(sel3(e, m)) <= eps * tst1) break;
            m++;
        }
        if (m > l) {
            int guard = 0;
            do {
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = snh_hypot(p, 1.0);
                double dl1, h, cc, c2, c3, el1, s, s2;
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                dl1 = d[l + 1];
                h = g - d[l];
#pragma unroll
                for (int i = l + 2; i < n; i++) d[i] -= h;
                f = f + h;
                p = sel3(d, m);
                cc = 1.0;
                c2 = cc;
                c3 = cc;
                el1 = e[l + 1];
                s = 0.0;
                s2 = 0.0;
                for (int i = m - 1; i >= l; i--) {
                    double ei = sel3(e, i), di = sel3(d, i);
                    c3 = c2;
                    c2 = cc;
                    s2 = s;
                    g = cc * ei;
                    h = cc * p;
                    r = snh_hypot(p, ei);
                    set3(e, i + 1, s * r);
                    s = ei / r;
                    cc = p / r;
                    p = cc * di - s * g;
                    set3(d, i + 1, h + s * (cc * g + s * di));
                    rot_cols(V, i, s, cc);
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = cc * p;
                /* the reference loop has no iteration cap; a cap of 64
                   sweeps (QL converges in a few) only guards the device */
            } while (snh_ref_iabs(e[l]) > eps * tst1 && ++guard < 64);
        }
        d[l] = d[l] + f;
        e[l] = 0.0;
    }
    /* ascending sort with eigenvector swaps */
#pragma unroll
    for (int i = 0; i < n - 1; i++) {
        double p = d[i];
        int k = i;
#pragma unroll
        for (int j = i + 1; j < n; j++)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            set3(d, k, d[i]);
            d[i] = p;
#pragma unroll
            for (int j = 0; j < n; j++) {
                double q = V[j][i];
                double vk = sel3(V[j], k);
                V[j][i] = vk;
                set3(V[j], k, q);
            }
        }
    }
}

__device__ __forceinline__ void principal_stresses(const double st[3][3], double d[3], double V[3][3]) {
    double e[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) V[i][j] = st[i][j];
    tred2_3(d, e, V);
    tql2_3(d, e, V);
}

/* piecewise-linear softening (plastic_BI/Constitutive.c:116-132,
   viscoplastic_BI/Constitutive.c:266-280): values persist when no segment
   matches */
__device__ __forceinline__ void soften(const DevMaterial& m, double ps, double& phi, double& psi, double& coh,
                                       double& hard) {
    for (uint32_t i = 0; i < m.nseg; i++) {
        const double pl1 = m.plstrain[i], pl2 = m.plstrain[i + 1];
        if (ps >= pl1 && ps <= pl2) {
            const double tgf = m.tgf[i];
            const double tgd = m.tgd[i];
            const double tgc = m.tgc[i];
            phi = m.fa[i] + tgf * (ps - pl1);
            psi = m.da[i] + tgd * (ps - pl1);
            coh = m.coh[i] + tgc * (ps - pl1);
            hard = tgc;
        }
    }
}

/* Mohr-Coulomb return in principal space and back-rotation
   (plastic_BI/Constitutive.c:159-223; viscoplastic_BI :307-371 with the
   hardening term).  st is the full 3x3 tensor; returns whether the tet
   yielded, with the plastic strain increment in dps. */
__device__ __forceinline__ bool mc_return(const DevMaterial& m, double st[3][3], double s[3], const double cn[3][3],
                                          double sphi, double spsi, double cohesion, double tension_cutoff,
                                          double hardening, bool with_hard, double& dps) {
    const double a1 = m.lambda + 2.0 * m.mu;
    const double a2 = m.lambda;
    double alam, dep1, dep2, dep3;
    double anphi = (1.0 + sphi) / (1.0 - sphi);
    double anpsi = (1.0 + spsi) / (1.0 - spsi);
    double fs = s[0] - s[2] * anphi + 2 * cohesion * sqrt(anphi);
    double ft = s[2] - tension_cutoff;
    if (!(fs < 0.0 || ft > 0.0)) return false;
    double aP = sqrt(1.0 + anphi * anphi) + anphi;
    double sP = tension_cutoff * anphi - 2 * cohesion * sqrt(anphi);
    double h = s[2] - tension_cutoff + aP * (s[0] - sP);
    if (h < 0.0) {
        if (with_hard)
            alam = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi + 2.0 * sqrt(anphi) * hardening);
        else
            alam = fs / (a1 - a2 * anpsi + a1 * anphi * anpsi - a2 * anphi);
        s[0] -= alam * (a1 - a2 * anpsi);
        s[1] -= alam * a2 * (1.0 - anpsi);
        s[2] -= alam * (a2 - a1 * anpsi);
        dep1 = alam;
        dep2 = 0.0;
        dep3 = -alam * anpsi;
    } else {
        alam = ft / a1;
        s[0] -= alam * a2;
        s[1] -= alam * a2;
        s[2] -= alam * a1;
        dep1 = 0.0;
        dep2 = 0.0;
        dep3 = alam;
    }
    double depm = snh_div3(dep1 + dep2 + dep3);
    dps = sqrt(0.5 * ((dep1 - depm) * (dep1 - depm) + (dep2 - depm) * (dep2 - depm) +
                      (dep3 - depm) * (dep3 - depm) + depm * depm));
#pragma unroll
    for (int mm = 0; mm < 3; mm++)
#pragma unroll
        for (int nn = 0; nn < 3; nn++) st[mm][nn] = 0.0;
#pragma unroll
    for (int mm = 0; mm < 3; mm++)
#pragma unroll
        for (int nn = mm; nn < 3; nn++)
#pragma unroll
            for (int k = 0; k < 3; k++) st[mm][nn] += cn[mm][k] * cn[nn][k] * s[k];
    return true;
}

/* sines/tangent of the softened angles: the host's libm values when the
   angles equal the first table entry (constant-angle tables), the device's
   otherwise */
__device__ __forceinline__ void angle_terms(const DevMaterial& m, double phi, double psi, double& sphi, double& spsi) {
    const double degrad = 0.017453292519943295; /* PI/180 */
    sphi = (phi == m.fa[0]) ? m.sphi : sin(phi * degrad);
    spsi = (psi == m.da[0]) ? m.spsi : sin(psi * degrad);
}
__device__ __forceinline__ double tan_phi(const DevMaterial& m, double phi) {
    const double degrad = 0.017453292519943295;
    return (phi == m.fa[0]) ? m.tanphi : tan(phi * degrad);
}

/* True when the trial stress st (00,11,22,01,02,12) cannot reach the
   Mohr-Coulomb or tension surface.  The principal stresses of a symmetric
   3x3 lie within p +- sqrt(2/3) |dev st| (p the mean stress; equality for
   an axisymmetric deviator), which bounds fs = s0 - s2 anphi + 2 c sqrt(anphi)
   from below and ft = s2 - tc from above.  The margin covers the eigen-
   solver's error, including the reference's int abs() that drops entries
   below 1 Pa.  For such a tet mc_return changes neither the stress nor the
   plastic strain, so its eigen-decomposition is skipped: the results are
   bit-identical. */
__device__ __forceinline__ bool mc_surely_elastic(const DevMaterial& m, const double st[6], double sphi, double cohesion,
                                                  double tc) {
    const double p = (st[0] + st[1] + st[2]) * (1.0 / 3.0);
    const double d0 = st[0] - p, d1 = st[1] - p, d2 = st[2] - p;
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * (st[3] * st[3] + st[4] * st[4] + st[5] * st[5]);
    const double r = sqrt(r2 * (2.0 / 3.0));
    const double anphi = sphi == m.sphi ? m.anphi0 : (1.0 + sphi) / (1.0 - sphi);
    const double smin = p - r, smax = p + r;
    const double scale = fabs(p) + r + fabs(cohesion) + fabs(tc);
    const double margin = 64.0 + 1e-9 * scale * (1.0 + anphi);
    const double fs_lo = smin - smax * anphi + 2.0 * cohesion * (sphi == m.sphi ? m.sqanphi0 : sqrt(anphi));
    const double ft_hi = smax - tc;
    return fs_lo > margin && ft_hi < -margin;
}

/* eigen-decomposition of the symmetric trial stress and Mohr-Coulomb return
   for a tet that may yield; out of line, so that the registers it needs are
   not reserved across the whole element loop (most tets never call it).
   io[0..5] holds the trial stress; on yield it holds the returned stress and
   io[6] the plastic strain increment. */
__device__ __noinline__ bool mc_plastic_update(const DevMaterial& m, double* io, double sphi, double spsi,
                                               double coh, double tc) {
    double S[3][3];
    S[0][0] = io[0]; S[1][1] = io[1]; S[2][2] = io[2];
    S[0][1] = S[1][0] = io[3];
    S[0][2] = S[2][0] = io[4];
    S[1][2] = S[2][1] = io[5];
    double d3[3], cn[3][3], dd = 0.0;
    principal_stresses(S, d3, cn);
    if (!mc_return(m, S, d3, cn, sphi, spsi, coh, tc, 0.0, false, dd)) return false;
    io[0] = S[0][0]; io[1] = S[1][1]; io[2] = S[2][2];
    io[3] = S[0][1]; io[4] = S[0][2]; io[5] = S[1][2];
    io[6] = dd;
    return true;
}

/* The call passes a local copy, made only on the path that calls, so the
   caller's st stays in registers instead of living in scratch memory for
   the whole kernel (an array whose address reaches an out-of-line call
   cannot be promoted). */
__device__ __forceinline__ bool mc_plastic_step(const DevMaterial& m, double st[6], double sphi, double spsi,
                                                double coh, double tc, double& dps) {
    double io[7] = {st[0], st[1], st[2], st[3], st[4], st[5], 0.0};
    if (!mc_plastic_update(m, io, sphi, spsi, coh, tc)) return false;
#pragma unroll
    for (int c = 0; c < 6; c++) st[c] = io[c];
    dps = io[6];
    return true;
}

/* viscosity of maxwell (:123-151) / viscoplastic_BI (:174-218) at the tet's
   mean nodal temperature avgTemp (zero without the temperature plugin) */
__device__ __forceinline__ double viscosity_of(const DevMaterial& m, double rvisc, double sd0, double sd1, double sd2,
                                               const double sn[6], double dt, double avgTemp = 0.0) {
    double visc;
    if (m.linear_visc) {
        visc = rvisc; /* pow(x, 0) * exp(0) == 1 exactly */
    } else {
        const double R = 8.31448;
        double srJ2 = sqrt(fabs(sd1 * sd2 + sd2 * sd0 + sd0 * sd1 - sn[3] * sn[3] - sn[4] * sn[4] - sn[5] * sn[5])) / dt;
        if (srJ2 == 0.0) srJ2 = m.refsrate;
        visc = rvisc * pow((srJ2 / m.refsrate), (1. / m.srexponent - 1.)) *
               exp(m.activationE / R * (1. / (avgTemp + 273.15) - 1. / (m.reftemp + 273.15)));
    }
    if (visc < m.vis_min) visc = m.vis_min;
    if (visc > m.vis_max) visc = m.vis_max;
    return visc;
}

/* temperature coupling of one tet with nodes nd[] (temperature plugin on):
   the mean nodal temperature of the viscous laws, 0.25 T summed in node
   order (maxwell/Constitutive.c:124-132, viscoplastic_BI :178-186), only
   needed by a non-linear law */
__device__ __forceinline__ double tet_temp(const StepArgs& a, const DevMaterial& m, const int nd[4]) {
    if (!a.temp || m.linear_visc) return 0.0;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) s += 0.25 * a.temp[nd[q]];
    return s;
}

/* thermal stress on the diagonal before the viscous update
   (maxwell/Constitutive.c:106-110, viscoplastic_BI :145-149) */
__device__ __forceinline__ void add_thermal_stress(const StepArgs& a, double st[6], int t, size_t e) {
    if (!a.tstress) return;
    const double ts = a.tstress[(size_t)t * (size_t)a.E + e];
    st[0] += ts;
    st[1] += ts;
    st[2] += ts;
}

/* the tet density set at the end of Stress (Stress.c:169-170) with beta = 0,
   from the tet average of the temperatures computed before the BCs
   (UpdateAverageTemp_LoopElements, temperature/Heat.c:135-160) */
__device__ __forceinline__ void store_density(const StepArgs& a, const DevMaterial& m, const int nd[4], int t,
                                              size_t e) {
    double avg = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) avg += a.tpre[nd[q]] / 4.0;
    a.dens[(size_t)t * (size_t)a.E + e] = m.phsDensity * (1.0 - m.alpha * (avg - m.reftemp));
}

/* deviatoric Maxwell update shared by maxwell and viscoplastic_BI */
__device__ __forceinline__ void maxwell_update(const DevMaterial& m, double st[6], const double sn[6],
                                               double trace_strain, double visc, double dt) {
    const double bulkm = m.lambda + snh_div3(2.0 * m.mu);
    const double rmu = m.mu;
    double trace_stress = st[0] + st[1] + st[2];
    const double ts3 = snh_div3(trace_stress), te3 = snh_div3(trace_strain);
    double straind0 = sn[0] - te3;
    double straind1 = sn[1] - te3;
    double straind2 = sn[2] - te3;
    double stressd0 = st[0] - ts3;
    double stressd1 = st[1] - ts3;
    double stressd2 = st[2] - ts3;
    double temp = rmu / (2.0 * visc) * dt;
    double vic1 = 1.0 - temp;
    double vic2 = 1.0 / (1.0 + temp);
    stressd0 = (stressd0 * vic1 + 2.0 * rmu * straind0) * vic2;
    stressd1 = (stressd1 * vic1 + 2.0 * rmu * straind1) * vic2;
    stressd2 = (stressd2 * vic1 + 2.0 * rmu * straind2) * vic2;
    st[3] = (st[3] * vic1 + 2.0 * rmu * sn[3]) * vic2;
    st[4] = (st[4] * vic1 + 2.0 * rmu * sn[4]) * vic2;
    st[5] = (st[5] * vic1 + 2.0 * rmu * sn[5]) * vic2;
    double VolumicStress = ts3 + bulkm * trace_strain;
    st[0] = stressd0 + VolumicStress;
    st[1] = stressd1 + VolumicStress;
    st[2] = stressd2 + VolumicStress;
}

/* deviatoric strains of a strain tensor for the viscosity law */
__device__ __forceinline__ void devstrain(const double sn[6], double trace_strain, double& d0, double& d1, double& d2) {
    const double te3 = snh_div3(trace_strain);
    d0 = sn[0] - te3;
    d1 = sn[1] - te3;
    d2 = sn[2] - te3;
}

/* --------------------------------------------------------------------------
 * k_elem: Snac_EP_CalcStresses with the pending rotation.
 *
 * One lane per tetrahedron: a wave holds 6 elements x 10 tets (lanes 60..63
 * idle).  Element-level sums (volume, strain-rate and stress trace averages)
 * are formed in tet order from LDS, so they round exactly as the reference's
 * sequential loops.
 * ------------------------------------------------------------------------ */
enum { RS_ELASTIC = 1, RS_PLASTIC = 2, RS_MAXWELL = 4, RS_VISCOPLASTIC = 8 };
#define EPW 6   /* elements per wave */

/* TetraToNode packed 3 bits per node, 12 bits per tet (TetrahedraTables.c:42-53) */
__host__ __device__ __forceinline__ int tet_node(int t, int q) {
    /* t in 0..9, q in 0..3: decode from two 64-bit words */
    constexpr unsigned long long P0 =
        (0x7ull << 0) | (0x2ull << 3) | (0x3ull << 6) | (0x0ull << 9) |          /* tet 0: 7 2 3 0 */
        (0x0ull << 12) | (0x5ull << 15) | (0x4ull << 18) | (0x7ull << 21) |      /* tet 1: 0 5 4 7 */
        (0x2ull << 24) | (0x5ull << 27) | (0x1ull << 30) | (0x0ull << 33) |      /* tet 2: 2 5 1 0 */
        (0x5ull << 36) | (0x2ull << 39) | (0x6ull << 42) | (0x7ull << 45) |      /* tet 3: 5 2 6 7 */
        (0x2ull << 48) | (0x5ull << 51) | (0x0ull << 54) | (0x7ull << 57);       /* tet 4: 2 5 0 7 */
    constexpr unsigned long long P1 =
        (0x3ull << 0) | (0x4ull << 3) | (0x7ull << 6) | (0x6ull << 9) |          /* tet 5: 3 4 7 6 */
        (0x4ull << 12) | (0x3ull << 15) | (0x0ull << 18) | (0x1ull << 21) |      /* tet 6: 4 3 0 1 */
        (0x6ull << 24) | (0x1ull << 27) | (0x2ull << 30) | (0x3ull << 33) |      /* tet 7: 6 1 2 3 */
        (0x1ull << 36) | (0x6ull << 39) | (0x5ull << 42) | (0x4ull << 45) |      /* tet 8: 1 6 5 4 */
        (0x1ull << 48) | (0x3ull << 51) | (0x6ull << 54) | (0x4ull << 57);       /* tet 9: 1 3 6 4 */
    const unsigned long long w = t < 5 ? P0 : P1;
    const int sh = (t < 5 ? t : t - 5) * 12 + q * 3;
    return (int)((w >> sh) & 7ull);
}

/* hex-local node -> node index (HexaEL.c:462-469 numbering) */
__device__ __forceinline__ int hex_node_index(const StepArgs& a, int ei, int ej, int ek, int q) {
    const int di = (q ^ (q >> 1)) & 1, dj = (q >> 1) & 1, dk = q >> 2;
    return (ei + di) + a.I * ((ej + dj) + a.J * (ek + dk));
}

/* tet_temp / store_density for tet t of element (ei, ej, ek) */
__device__ __forceinline__ double tet_temp_at(const StepArgs& a, const DevMaterial& m, int ei, int ej, int ek, int t) {
    if (!a.temp || m.linear_visc) return 0.0;
    int nd[4];
#pragma unroll
    for (int q = 0; q < 4; q++) nd[q] = hex_node_index(a, ei, ej, ek, kTetraToNode[t][q]);
    return tet_temp(a, m, nd);
}

__device__ __forceinline__ void store_density_at(const StepArgs& a, const DevMaterial& m, int ei, int ej, int ek,
                                                 int t, size_t e) {
    int nd[4];
#pragma unroll
    for (int q = 0; q < 4; q++) nd[q] = hex_node_index(a, ei, ej, ek, kTetraToNode[t][q]);
    store_density(a, m, nd, t, e);
}

/* softened parameters carried along the tets of an element: the reference
   declares them once per element (plastic_BI/Constitutive.c:78-81,
   viscoplastic_BI/Constitutive.c:115-119) */
struct Soft {
    double phi, psi, coh, hard;
};

/* TW = false: one 64-lane block holds 6 elements x 10 tets.
   TW = true ("tet per wave"): a 640-thread block holds 64 elements, wave w
   taking tet w of all of them, so every stress and node load of a wave is
   over 64 consecutive elements. */
template <int RSET, bool TW, bool MM>
__device__ __forceinline__ void elem_pass(StepArgs a, const int blk, const int nblk) {
    constexpr int NE = TW ? 64 : EPW;       /* elements per block */
    constexpr int NT = TW ? NTET + 1 : NTET; /* padded against LDS bank conflicts */
    __shared__ double sh_a[NE][NT];
    __shared__ double sh_b[NE][NT];
    __shared__ double sh_c[NE][NT];   /* the stress-trace sums */
    __shared__ double sh_sp[NE][NT][5];   /* matched, phi, psi, coh, hard */
    __shared__ double sh_ps[NE][NT];
    const int lane = threadIdx.x;
    /* (tet-per-lane) lanes 60..63 carry no tet: they mirror tet 0 of element
       slot 0 but never write LDS or memory */
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
    const double dt_rot = a.scal->dt_rot;
    const DevMaterial& m = mat_of<MM>(a, ec);

    /* With the Maxwell law alone, the tet's stored stress and the element's
       viscosity are loaded ahead of the kinematics and the block barrier, so
       that their latency overlaps them (3 % off the pass; with the other
       laws the same order measured 1-3 % slower). */
    constexpr bool kEarly = RSET == RS_MAXWELL;
    double st[6];
    double rvisc_e = 0.0;
    if constexpr (kEarly) {
#pragma unroll
        for (int c = 0; c < 6; c++) st[c] = active ? a.stress[((size_t)c * NTET + t) * E + ec] : 0.0;
        rvisc_e = a.refvisc ? a.refvisc[ec] : m.refvisc;
    }
    /* the plastic laws' stored plastic strain, likewise ahead of the
       kinematics (1.6 % off the Mohr-Coulomb pass) */
    const bool is_pl = (RSET & RS_PLASTIC) && (m.rheology & SNH_RHEOLOGY_PLASTIC);
    const bool is_vp = (RSET & RS_VISCOPLASTIC) && (m.rheology & SNH_RHEOLOGY_VISCOPLASTIC);
    double ps = 0.0;
    if ((RSET & (RS_PLASTIC | RS_VISCOPLASTIC)) && (is_pl || is_vp) && active) ps = a.ps[(size_t)t * E + ec];

    /* kinematics of this lane's tet (UpdateElement.c:84-219, StrainRate.c:65-184) */
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
    const double trace = snh_div3(k.sr[0] + k.sr[1] + k.sr[2]);
    if (in_wave) {
        sh_a[le][t] = k.V;
        sh_b[le][t] = trace * k.V;
    }
    __syncthreads();
    double evol = 0.0, partial = 0.0;
#pragma unroll
    for (int q = 0; q < NTET; q++) {
        evol += sh_a[le][q];        /* UpdateElement.c:141 */
        partial += sh_b[le][q];     /* StrainRate.c:281 */
    }
    evol *= 0.5;                          /* UpdateElement.c:267-268 */
    const double elemVolume = 2.0 * evol; /* StrainRate.c:57, Stress.c:68 */
    const double srcorr = partial / elemVolume;

    /* rotation of the stored stress, strain (Stress.c:71-82) */
    if constexpr (!kEarly) {
#pragma unroll
        for (int c = 0; c < 6; c++) st[c] = active ? a.stress[((size_t)c * NTET + t) * E + ec] : 0.0;
    }
    rotate_stress(st, k.w, dt_rot);
    double sn[6];
#pragma unroll
    for (int c = 0; c < 3; c++) sn[c] = (k.sr[c] + (-1.0 * trace + srcorr)) * dt;
#pragma unroll
    for (int c = 3; c < 6; c++) sn[c] = k.sr[c] * dt;

    Soft sp = {0.0, 0.0, 0.0, 0.0};
    if (RSET & (RS_PLASTIC | RS_VISCOPLASTIC)) {
        /* own segment match, then resolve the carried parameters in tet order */
        double matched = 0.0;
        Soft own = {0.0, 0.0, 0.0, 0.0};
        if (is_pl || is_vp) {
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
        if ((is_pl || is_vp) && matched != 0.0 && own.phi >= 0.0) {
            /* the scan below ends with this tet's own parameters whenever its
               plastic strain lies in a segment and the angle is not negative */
            sp = own;
        } else if (is_pl || is_vp) {
            sp.phi = is_pl ? 30.0 : 0.0;   /* plastic_BI :78-80 defaults */
            sp.coh = is_pl ? 30.0e6 : 0.0;
            for (int q = 0; q <= t; q++) {
                if (sh_sp[le][q][0] != 0.0) {
                    sp.phi = sh_sp[le][q][1];
                    sp.psi = sh_sp[le][q][2];
                    sp.coh = sh_sp[le][q][3];
                    sp.hard = sh_sp[le][q][4];
                }
                if (!(sp.phi >= 0.0) && sh_ps[le][q] < 0.0) {
                    /* negative plastic strain is reset (plastic_BI :143-151) */
                    sp.phi = m.fa[0];
                    sp.psi = m.da[0];
                    sp.coh = m.coh[0];
                    if (q == t) ps = 0.0;
                }
            }
        }
    }

    if (is_pl) {
        double trace_strain = sn[0] + sn[1] + sn[2];
        double S[3][3];
        st[0] += (2.0 * m.mu) * sn[0] + m.lambda * (trace_strain);
        st[1] += (2.0 * m.mu) * sn[1] + m.lambda * (trace_strain);
        st[2] += (2.0 * m.mu) * sn[2] + m.lambda * (trace_strain);
        st[3] += (2.0 * m.mu) * sn[3];
        st[4] += (2.0 * m.mu) * sn[4];
        st[5] += (2.0 * m.mu) * sn[5];
        S[0][0] = st[0]; S[1][1] = st[1]; S[2][2] = st[2];
        S[0][1] = S[1][0] = st[3];
        S[0][2] = S[2][0] = st[4];
        S[1][2] = S[2][1] = st[5];
        const double coh = a.cohoff ? sp.coh + a.cohoff[ec] : sp.coh;
        double tc = m.ten_off;
        if (sp.phi > 0.0 && tc == 0.0) tc = coh / tan_phi(m, sp.phi);
        double sphi, spsi, dps = 0.0;
        angle_terms(m, sp.phi, sp.psi, sphi, spsi);
        bool changed = false;
        if (!mc_surely_elastic(m, st, sphi, coh, tc) && mc_plastic_step(m, st, sphi, spsi, coh, tc, dps)) {
            ps = ps + dps;
            changed = true;
        }
        /* in place, an unchanged value need not be stored */
        if (active && (changed || a.ps_out != a.ps || ps != sh_ps[le][t])) a.ps_out[(size_t)t * E + ec] = ps;
    } else if (is_vp) {
        add_thermal_stress(a, st, t, ec);
        double trace_strain = sn[0] + sn[1] + sn[2];
        double d0, d1, d2;
        devstrain(sn, trace_strain, d0, d1, d2);
        double rvisc = kEarly ? rvisc_e : (a.refvisc ? a.refvisc[ec] : m.refvisc);
        double visc = viscosity_of(m, rvisc, d0, d1, d2, sn, dt, tet_temp(a, m, nd));
        if (a.visc && active) a.visc[(size_t)t * E + ec] = visc;
        maxwell_update(m, st, sn, trace_strain, visc, dt);
        /* unsymmetrised: the lower triangle is zero (viscoplastic_BI :246) */
        double S[3][3] = {{st[0], st[3], st[4]}, {0.0, st[1], st[5]}, {0.0, 0.0, st[2]}};
        double d3[3], cn[3][3];
        principal_stresses(S, d3, cn);
        const double coh = a.cohoff ? sp.coh + a.cohoff[ec] : sp.coh;
        double tc = m.ten_off;
        if (sp.phi > 0.0 && tc == 0.0) tc = coh / tan_phi(m, sp.phi);
        double sphi, spsi, dps = 0.0;
        angle_terms(m, sp.phi, sp.psi, sphi, spsi);
        if (mc_return(m, S, d3, cn, sphi, spsi, coh, tc, sp.hard, true, dps)) {
            ps = ps + dps;
            st[0] = S[0][0]; st[1] = S[1][1]; st[2] = S[2][2];
            st[3] = S[0][1]; st[4] = S[0][2]; st[5] = S[1][2];
        }
        if (active) a.ps_out[(size_t)t * E + ec] = ps;
    } else if ((RSET & RS_MAXWELL) && (m.rheology & SNH_RHEOLOGY_MAXWELL)) {
        /* trace strain from the volume ratio (maxwell/Constitutive.c:113) */
        double Xp[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            Xp[q][0] = a.xp[nd[q]];
            Xp[q][1] = a.xp[a.N + nd[q]];
            Xp[q][2] = a.xp[2 * a.N + nd[q]];
        }
        double old_volume = tet_volume(Xp[0], Xp[1], Xp[2], Xp[3]);
        double trace_strain = k.V / old_volume - 1.0;
        add_thermal_stress(a, st, t, ec);
        double d0, d1, d2;
        devstrain(sn, trace_strain, d0, d1, d2);
        double rvisc = kEarly ? rvisc_e : (a.refvisc ? a.refvisc[ec] : m.refvisc);
        double visc = viscosity_of(m, rvisc, d0, d1, d2, sn, dt, tet_temp(a, m, nd));
        if (a.visc && active) a.visc[(size_t)t * E + ec] = visc;
        maxwell_update(m, st, sn, trace_strain, visc, dt);
    } else if ((RSET & RS_ELASTIC) && (m.rheology & SNH_RHEOLOGY_ELASTIC)) {
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

    /* stress trace averaging (Stress.c:87-137) */
    const double tr = snh_div3(st[0] + st[1] + st[2]);
    /* its own array: no wait for the readers of sh_b */
    if (in_wave) sh_c[le][t] = tr * k.V;
    __syncthreads();
    double pst = 0.0;
#pragma unroll
    for (int q = 0; q < NTET; q++) pst += sh_c[le][q];
    const double avgTrace = pst / elemVolume;
    st[0] += -1.0 * tr + avgTrace;
    st[1] += -1.0 * tr + avgTrace;
    st[2] += -1.0 * tr + avgTrace;
    if (active) {
#pragma unroll
        for (int c = 0; c < 6; c++) a.stress_out[((size_t)c * NTET + t) * E + ec] = st[c];
        if (a.tpre) store_density(a, m, nd, t, ec);
    }
}

template <int RSET, bool TW, bool MM>
__global__ void __launch_bounds__(TW ? 640 : 64) k_elem(StepArgs a) {
    elem_pass<RSET, TW, MM>(a, (int)blockIdx.x, (int)gridDim.x);
}

/* --------------------------------------------------------------------------
 * k_elem_seq: the same element pass with one thread per element, the 10 tets
 * in sequence (kept as a second variant: for laws dominated by per-tet
 * iterative work, e.g. the plastic eigen-solve, one element per lane
 * diverges less than one tet per lane).  Identical arithmetic.
 * ------------------------------------------------------------------------ */
template <int RSET, bool LOWREG, bool MM>
__global__ void __launch_bounds__(ELEM_BLOCK) k_elem_seq(StepArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int ei, ej, ek;
    elem_ijk(a, e, ei, ej, ek);
    const double dt = a.scal->dt;
    const double dt_rot = a.scal->dt_rot;
    const DevMaterial& m = mat_of<MM>(a, e);

    double X[8][3], U[8][3];
    load_hex(a, a.xc, ei, ej, ek, X);
    load_hex(a, a.v, ei, ej, ek, U);

    /* pass 1: kinematics of the 10 tets.  LOWREG keeps none of it and
       recomputes it in pass 2 (bit-identical), trading FP64 work for
       registers */
    double sr[LOWREG ? 1 : NTET][6], w[LOWREG ? 1 : NTET][3], vol[LOWREG ? 1 : NTET];
    double trace[LOWREG ? 1 : NTET];
    double evol = 0.0, partial = 0.0;
#pragma unroll
    for (int t = 0; t < NTET; t++) {
        double Xt[4][3], Ut[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                Xt[q][d] = X[kTetraToNode[t][q]][d];
                Ut[q][d] = U[kTetraToNode[t][q]][d];
            }
        TetKin k;
        tet_kinematics(Xt, Ut, k);
        evol += k.V;
        const double trc = snh_div3(k.sr[0] + k.sr[1] + k.sr[2]);
        partial += trc * k.V;
        if constexpr (!LOWREG) {
            vol[t] = k.V;
            trace[t] = trc;
#pragma unroll
            for (int c = 0; c < 6; c++) sr[t][c] = k.sr[c];
#pragma unroll
            for (int c = 0; c < 3; c++) w[t][c] = k.w[c];
        }
    }
    evol *= 0.5;                        /* UpdateElement.c:267-268 */
    const double elemVolume = 2.0 * evol; /* StrainRate.c:57, Stress.c:68 */
    const double srcorr = partial / elemVolume;
    if constexpr (!LOWREG) {
        if (a.diag) {
            /* element strain-rate invariant (StrainRate.c:272-304) */
            double srVolAvg = 0.0, srOtherAvg = 0.0;
#pragma unroll
            for (int t = 0; t < NTET; t++) {
                const double corr = -1.0 * trace[t] + partial / elemVolume;
                const double s00 = sr[t][0] + corr, s11 = sr[t][1] + corr, s22 = sr[t][2] + corr;
                srVolAvg += s11 * s22 + s22 * s00 + s00 * s11;
                srOtherAvg += sr[t][3] * sr[t][3] + sr[t][5] * sr[t][5] + sr[t][4] * sr[t][4];
            }
            srVolAvg /= (double)NTET;
            srOtherAvg /= (double)NTET;
            a.diag[e] = 0.5 * sqrt(0.5 * fabs(-1.0 * srVolAvg + srOtherAvg));
        }
    }

    /* pass 2: rotation, strain, constitutive */
    double sig[LOWREG ? 1 : NTET][6];
    double trs[LOWREG ? NTET : 1], pst = 0.0;
    const size_t E = (size_t)a.E;
    /* plastic_BI / viscoplastic_BI keep their softened parameters across the
       tets of an element */
    double pcoh = (RSET & RS_PLASTIC) && (m.rheology & SNH_RHEOLOGY_PLASTIC) ? 30.0e6 : 0.0;
    double pphi = (RSET & RS_PLASTIC) && (m.rheology & SNH_RHEOLOGY_PLASTIC) ? 30.0 : 0.0;
    double ppsi = 0.0, phard = 0.0, ptc = 0.0;
    const double cohoff = a.cohoff ? a.cohoff[e] : 0.0;
    double depls = 0.0, totalVolume = 0.0;   /* aps (plastic_BI/Constitutive.c:232-236) */
#pragma unroll
    for (int t = 0; t < NTET; t++) {
        TetKin kr;
        const double *srt, *wt;
        double volt, tracet;
        if constexpr (LOWREG) {
            double Xt[4][3], Ut[4][3];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    Xt[q][d] = X[kTetraToNode[t][q]][d];
                    Ut[q][d] = U[kTetraToNode[t][q]][d];
                }
            tet_kinematics(Xt, Ut, kr);
            srt = kr.sr;
            wt = kr.w;
            volt = kr.V;
            tracet = snh_div3(kr.sr[0] + kr.sr[1] + kr.sr[2]);
        } else {
            srt = sr[t];
            wt = w[t];
            volt = vol[t];
            tracet = trace[t];
        }
        double st[6];
#pragma unroll
        for (int c = 0; c < 6; c++) st[c] = a.stress[((size_t)c * NTET + t) * E + e];
        rotate_stress(st, wt, dt_rot);
        double s6[6];
#pragma unroll
        for (int c = 0; c < 3; c++) s6[c] = srt[c] + (-1.0 * tracet + srcorr);
#pragma unroll
        for (int c = 3; c < 6; c++) s6[c] = srt[c];
        double sn[6];
#pragma unroll
        for (int c = 0; c < 6; c++) sn[c] = s6[c] * dt;

        if ((RSET & RS_PLASTIC) && (m.rheology & SNH_RHEOLOGY_PLASTIC)) {
            double plasticStrain = a.ps[(size_t)t * E + e];
            double trace_strain = sn[0] + sn[1] + sn[2];
            double S[3][3];
            st[0] += (2.0 * m.mu) * sn[0] + m.lambda * (trace_strain);
            st[1] += (2.0 * m.mu) * sn[1] + m.lambda * (trace_strain);
            st[2] += (2.0 * m.mu) * sn[2] + m.lambda * (trace_strain);
            st[3] += (2.0 * m.mu) * sn[3];
            st[4] += (2.0 * m.mu) * sn[4];
            st[5] += (2.0 * m.mu) * sn[5];
            S[0][0] = st[0]; S[1][1] = st[1]; S[2][2] = st[2];
            S[0][1] = S[1][0] = st[3];
            S[0][2] = S[2][0] = st[4];
            S[1][2] = S[2][1] = st[5];
            soften(m, plasticStrain, pphi, ppsi, pcoh, phard);
            double coh = a.cohoff ? pcoh + cohoff : pcoh;
            if (pphi >= 0.0) {
                ptc = m.ten_off;
                if (pphi > 0.0 && ptc == 0.0) ptc = coh / tan_phi(m, pphi);
            } else {
                /* negative plastic strain is reset (plastic_BI :143-151); a
                   negative friction angle otherwise is an error the host
                   checks before the run */
                plasticStrain = 0.0;
                pphi = m.fa[0];
                ppsi = m.da[0];
                pcoh = m.coh[0];
                coh = a.cohoff ? pcoh + cohoff : pcoh;
            }
            double sphi, spsi;
            angle_terms(m, pphi, ppsi, sphi, spsi);
            double dps = 0.0;
            if (!mc_surely_elastic(m, st, sphi, coh, ptc) && mc_plastic_step(m, st, sphi, spsi, coh, ptc, dps))
                plasticStrain = plasticStrain + dps;
            /* in place, an unchanged value need not be stored */
            if (a.ps_out != a.ps || plasticStrain != a.ps[(size_t)t * E + e])
                a.ps_out[(size_t)t * E + e] = plasticStrain;
            depls += plasticStrain * volt;
            totalVolume += volt;
        } else if ((RSET & RS_VISCOPLASTIC) && (m.rheology & SNH_RHEOLOGY_VISCOPLASTIC)) {
            double plasticStrain = a.ps[(size_t)t * E + e];
            add_thermal_stress(a, st, t, e);
            double trace_strain = sn[0] + sn[1] + sn[2];
            double d0, d1, d2;
            devstrain(sn, trace_strain, d0, d1, d2);
            double rvisc = a.refvisc ? a.refvisc[e] : m.refvisc;
            double visc = viscosity_of(m, rvisc, d0, d1, d2, sn, dt, tet_temp_at(a, m, ei, ej, ek, t));
            if (a.visc) a.visc[(size_t)t * E + e] = visc;
            maxwell_update(m, st, sn, trace_strain, visc, dt);
            /* unsymmetrised: the lower triangle is zero (viscoplastic_BI :246) */
            double S[3][3] = {{st[0], st[3], st[4]}, {0.0, st[1], st[5]}, {0.0, 0.0, st[2]}};
            double d3[3], cn[3][3];
            principal_stresses(S, d3, cn);
            soften(m, plasticStrain, pphi, ppsi, pcoh, phard);
            double coh = a.cohoff ? pcoh + cohoff : pcoh;
            if (pphi >= 0.0) {
                ptc = m.ten_off;
                if (pphi > 0.0 && ptc == 0.0) ptc = coh / tan_phi(m, pphi);
            } else {
                plasticStrain = 0.0;
                pphi = m.fa[0];
                ppsi = m.da[0];
                pcoh = m.coh[0];
                coh = a.cohoff ? pcoh + cohoff : pcoh;
            }
            double sphi, spsi;
            angle_terms(m, pphi, ppsi, sphi, spsi);
            double dps = 0.0;
            if (mc_return(m, S, d3, cn, sphi, spsi, coh, ptc, phard, true, dps)) {
                plasticStrain = plasticStrain + dps;
                st[0] = S[0][0]; st[1] = S[1][1]; st[2] = S[2][2];
                st[3] = S[0][1]; st[4] = S[0][2]; st[5] = S[1][2];
            }
            /* in place, an unchanged value need not be stored */
            if (a.ps_out != a.ps || plasticStrain != a.ps[(size_t)t * E + e])
                a.ps_out[(size_t)t * E + e] = plasticStrain;
            depls += plasticStrain * volt;
            totalVolume += volt;
        } else if ((RSET & RS_MAXWELL) && (m.rheology & SNH_RHEOLOGY_MAXWELL)) {
            /* trace strain from the volume ratio (maxwell/Constitutive.c:113) */
            double Xp[4][3];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int nd = node_index(a, ei + kHexOff[kTetraToNode[t][q]][0], ej + kHexOff[kTetraToNode[t][q]][1],
                                    ek + kHexOff[kTetraToNode[t][q]][2]);
                Xp[q][0] = a.xp[nd];
                Xp[q][1] = a.xp[a.N + nd];
                Xp[q][2] = a.xp[2 * a.N + nd];
            }
            double old_volume = tet_volume(Xp[0], Xp[1], Xp[2], Xp[3]);
            double trace_strain = volt / old_volume - 1.0;
            add_thermal_stress(a, st, t, e);
            double d0, d1, d2;
            devstrain(sn, trace_strain, d0, d1, d2);
            double rvisc = a.refvisc ? a.refvisc[e] : m.refvisc;
            double visc = viscosity_of(m, rvisc, d0, d1, d2, sn, dt, tet_temp_at(a, m, ei, ej, ek, t));
            if (a.visc) a.visc[(size_t)t * E + e] = visc;
            maxwell_update(m, st, sn, trace_strain, visc, dt);
        } else if ((RSET & RS_ELASTIC) && (m.rheology & SNH_RHEOLOGY_ELASTIC)) {
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
        if constexpr (LOWREG) {
            /* store now; pass 3 adds the averaged trace to the diagonal */
            trs[t] = snh_div3(st[0] + st[1] + st[2]);
            pst += trs[t] * volt;
#pragma unroll
            for (int c = 0; c < 6; c++) a.stress_out[((size_t)c * NTET + t) * E + e] = st[c];
        } else {
#pragma unroll
            for (int c = 0; c < 6; c++) sig[t][c] = st[c];
        }
    }

    /* pass 3: stress trace averaging (Stress.c:87-137) */
    if constexpr (LOWREG) {
        const double avgTrace = pst / elemVolume;
#pragma unroll
        for (int t = 0; t < NTET; t++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double* p = &a.stress_out[((size_t)c * NTET + t) * E + e];
                *p = *p + (-1.0 * trs[t] + avgTrace);
            }
    } else {
        double tr[NTET];
#pragma unroll
        for (int t = 0; t < NTET; t++) {
            tr[t] = snh_div3(sig[t][0] + sig[t][1] + sig[t][2]);
            pst += tr[t] * vol[t];
        }
        const double avgTrace = pst / elemVolume;
        double sVolAvg = 0.0, sOtherAvg = 0.0, pressure = 0.0;
#pragma unroll
        for (int t = 0; t < NTET; t++) {
            sig[t][0] += -1.0 * tr[t] + avgTrace;
            sig[t][1] += -1.0 * tr[t] + avgTrace;
            sig[t][2] += -1.0 * tr[t] + avgTrace;
#pragma unroll
            for (int c = 0; c < 6; c++) a.stress_out[((size_t)c * NTET + t) * E + e] = sig[t][c];
            if (a.diag) {
                /* element stress invariant and pressure (Stress.c:140-156) */
                const double s00 = sig[t][0] - avgTrace, s11 = sig[t][1] - avgTrace, s22 = sig[t][2] - avgTrace;
                sVolAvg += s11 * s22 + s22 * s00 + s00 * s11;
                sOtherAvg += sig[t][3] * sig[t][3] + sig[t][5] * sig[t][5] + sig[t][4] * sig[t][4];
                pressure += snh_div3(sig[t][0] + sig[t][1] + sig[t][2]);
            }
        }
        if (a.diag) {
            sVolAvg /= (double)NTET;
            sOtherAvg /= (double)NTET;
            pressure /= (double)NTET;
            a.diag[E + e] = 0.5 * sqrt(0.5 * fabs(-1.0 * sVolAvg + sOtherAvg));
            a.diag[2 * E + e] = -1.0 * pressure;
            a.diag[3 * E + e] = totalVolume > 0.0 ? depls / totalVolume : 0.0;
        }
    }
    if (a.tpre)
        for (int t = 0; t < NTET; t++) store_density_at(a, m, ei, ej, ek, t, e);
}

/* --------------------------------------------------------------------------
 * k_rotated: stress with the pending rotation applied, for downloads
 * ------------------------------------------------------------------------ */
__global__ void __launch_bounds__(ELEM_BLOCK) k_rotated(StepArgs a, double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int ei, ej, ek;
    elem_ijk(a, e, ei, ej, ek);
    const double dt_rot = a.scal->dt_rot;
    double X[8][3], U[8][3];
    load_hex(a, a.xc, ei, ej, ek, X);
    load_hex(a, a.v, ei, ej, ek, U);
    const size_t E = (size_t)a.E;
#pragma unroll
    for (int t = 0; t < NTET; t++) {
        double Xt[4][3], Ut[4][3];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                Xt[q][d] = X[kTetraToNode[t][q]][d];
                Ut[q][d] = U[kTetraToNode[t][q]][d];
            }
        TetKin k;
        tet_kinematics(Xt, Ut, k);
        double st[6];
#pragma unroll
        for (int c = 0; c < 6; c++) st[c] = a.stress[((size_t)c * NTET + t) * E + e];
        rotate_stress(st, k.w, dt_rot);
#pragma unroll
        for (int c = 0; c < 6; c++) out[((size_t)c * NTET + t) * E + e] = st[c];
    }
}

/* volumes of the current geometry, for downloads */
__global__ void k_volumes(StepArgs a, double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int ei, ej, ek;
    elem_ijk(a, e, ei, ej, ek);
    double X[8][3];
    load_hex(a, a.xc, ei, ej, ek, X);
#pragma unroll
    for (int t = 0; t < NTET; t++)
        out[(size_t)t * a.E + e] = tet_volume(X[kTetraToNode[t][0]], X[kTetraToNode[t][1]], X[kTetraToNode[t][2]],
                                              X[kTetraToNode[t][3]]);
}

/* --------------------------------------------------------------------------
 * k_node: Snac_EP_LoopNodesMomentum
 * ------------------------------------------------------------------------ */
#define GETSIGN(val) ((val) < 0 ? -1.0 : (val) > 0 ? +1.0 : 0.0)

/* the tet densities of Snac_Stress with the pressure-dependent term
   (Stress.c:147-170), for contexts with a material whose beta != 0: the
   element pressure is the mean over its tets of the mean of each final tet
   stress, and density = phsDensity (1 - alpha (avgTemp - reftemp) + beta
   pressure), avgTemp the tet's mean nodal temperature before the BCs with
   the temperature plugin, 0 without.  Runs after the element pass. */
template <bool MM, bool TH>
__global__ void __launch_bounds__(GEO_BLOCK) k_density(StepArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    const size_t E = (size_t)a.E;
    double pressure = 0.0;
#pragma unroll
    for (int t = 0; t < NTET; t++)
        pressure += snh_div3(a.stress_out[((size_t)0 * NTET + t) * E + e] + a.stress_out[((size_t)1 * NTET + t) * E + e] +
                             a.stress_out[((size_t)2 * NTET + t) * E + e]);
    pressure /= (double)NTET;
    const DevMaterial& m = mat_of<MM>(a, e);
    int ei, ej, ek;
    elem_ijk(a, e, ei, ej, ek);
#pragma unroll
    for (int t = 0; t < NTET; t++) {
        double avg = 0.0;
        if constexpr (TH) {
#pragma unroll
            for (int q = 0; q < 4; q++) avg += a.tpre[hex_node_index(a, ei, ej, ek, kTetraToNode[t][q])] / 4.0;
        }
        a.dens[(size_t)t * E + e] = m.phsDensity * (1.0 - m.alpha * (avg - m.reftemp) + m.beta * pressure);
    }
}

/* node handled by gather thread g.  Rows of I nodes are taken in bands of
   node_band k-planes, j outer and k inner within a band: the four node rows
   that read one element row's stress ((j,k), (j+1,k), (j,k+1), (j+1,k+1))
   then run close together in time on one XCD and share its L2, where the
   plain order puts a whole k-plane (J rows) between them.  node_band = 0
   keeps the plain order.  A bijection of [0, N). */
__device__ __forceinline__ int gather_node(const StepArgs& a, int g, int& i, int& j, int& k) {
    const int KB = a.node_band;
    if (KB <= 0) {
        node_ijk(a, g, i, j, k);
        return g;
    }
    const int r = (int)fastdiv((uint32_t)g, a.dI);
    i = g - r * a.I;
    const int per = a.J * KB;
    const int nfull = (int)fastdiv((uint32_t)a.K, a.dKB);
    if (r < nfull * per) {
        const int b = (int)fastdiv((uint32_t)r, a.dPer), rem = r - b * per;
        j = (int)fastdiv((uint32_t)rem, a.dKB);
        k = b * KB + (rem - j * KB);
    } else {
        const int kl = a.K - nfull * KB;
        const int rem = r - nfull * per;
        j = rem / kl;
        k = nfull * KB + (rem - j * kl);
    }
    return i + a.I * (j + a.J * k);
}

/* a node's own state for its update: coordinates, velocity and velocity
   condition, loaded ahead of its gather so that their latency overlaps it */
struct NodeOwn {
    double x0, x1, x2, v0, v1, v2;
    uint32_t code, mask;
    double b0, b1, b2;
};
__device__ __forceinline__ NodeOwn node_own(const StepArgs& a, int n) {
    const size_t N = (size_t)a.N;
    NodeOwn s;
    s.x0 = a.xc[n];
    s.x1 = a.xc[N + n];
    s.x2 = a.xc[2 * N + n];
    s.v0 = a.v[n];
    s.v1 = a.v[N + n];
    s.v2 = a.v[2 * N + n];
    s.code = a.bc[n];
    s.mask = 0u;
    s.b0 = s.b1 = s.b2 = 0.0;
    if (s.code) {
        const DevBC& b = a.bcs[s.code];
        s.mask = b.mask;
        s.b0 = b.val[0];
        s.b1 = b.val[1];
        s.b2 = b.val[2];
    }
    return s;
}

/* the second node pass for node n (UpdateNode.c:60-158) once its force f
   and inertial mass im are gathered: damping, velocity, velocity BC and
   coordinates (x: the node's current coordinates); with a replay tape, the
   undamped force, the mass and the damping branch.  Returns |v| for a
   constrained node, -inf otherwise (the Sync's vmax, Context.c:1218-1235). */
__device__ __forceinline__ double node_finish(const StepArgs& a, int n, double f0, double f1, double f2, double im,
                                              const NodeOwn& own, double dt) {
    double vmag = -INFINITY;
    const size_t N = (size_t)a.N;
    const double x0 = own.x0, x1 = own.x1, x2 = own.x2;
    double v0 = own.v0, v1 = own.v1, v2 = own.v2;
    if (a.tape_f) {
        /* what the reverse pass needs: the undamped force and the mass */
        a.tape_f[n] = f0;
        a.tape_f[N + n] = f1;
        a.tape_f[2 * N + n] = f2;
        a.tape_m[n] = im;
    }
    if (a.tape_code) {
        /* the damping branch of each component (the adjoint's tape, or a
           recorded schedule) */
        const int c0 = (int)(GETSIGN(v0) * GETSIGN(f0)) + 1;
        const int c1 = (int)(GETSIGN(v1) * GETSIGN(f1)) + 1;
        const int c2 = (int)(GETSIGN(v2) * GETSIGN(f2)) + 1;
        a.tape_code[n] = (uint8_t)(c0 | (c1 << 2) | (c2 << 4));
    }
    double sv0 = GETSIGN(v0), sv1 = GETSIGN(v1), sv2 = GETSIGN(v2);
    if (a.damp_sched) {
        const uint8_t ds = a.damp_sched[n];
        sv0 = (double)((int)(ds & 3u) - 1) * GETSIGN(f0);
        sv1 = (double)((int)((ds >> 2) & 3u) - 1) * GETSIGN(f1);
        sv2 = (double)((int)((ds >> 4) & 3u) - 1) * GETSIGN(f2);
    }
    f0 -= a.demf * sv0 * fabs(f0);
    f1 -= a.demf * sv1 * fabs(f1);
    f2 -= a.demf * sv2 * fabs(f2);
    if (a.force_out) {
        a.force_out[n] = f0;
        a.force_out[N + n] = f1;
        a.force_out[2 * N + n] = f2;
    }
    v0 += dt * f0 / im;
    v1 += dt * f1 / im;
    v2 += dt * f2 / im;
    if (own.code) {
        if (own.mask & 1u) v0 = own.b0;
        if (own.mask & 2u) v1 = own.b1;
        if (own.mask & 4u) v2 = own.b2;
        vmag = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    }
    /* a NaN or Inf in stress, mass or force reaches the velocities within
       the step: the counterpart of the reference's release-mode firewall
       on the element stress (Stress.c:161-166) */
    if (!(isfinite(v0) && isfinite(v1) && isfinite(v2))) a.scal->nonfinite = 1;
    a.v_out[n] = v0;
    a.v_out[N + n] = v1;
    a.v_out[2 * N + n] = v2;
    a.xn[n] = x0 + dt * v0;
    a.xn[N + n] = x1 + dt * v1;
    a.xn[2 * N + n] = x2 + dt * v2;
    return vmag;
}

template <bool MM, bool TH>
__device__ __forceinline__ void node_pass(StepArgs a, const int blk, const int nblk) {
    __shared__ double sh[NODE_BLOCK / 64];
    const int g = xcd_block(blk, nblk) * blockDim.x + threadIdx.x;
    double vmag = -INFINITY;
    if (g < a.N) {
        int i, j, k;
        const int n = gather_node(a, g, i, j, k);
        const double dt = a.scal->dt;
        const double c = a.scal->speedOfSound;
        const size_t E = (size_t)a.E;
        const NodeOwn own = node_own(a, n);
        double im = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
        for (int slot = 0; slot < 8; slot++) {
            const int ei = i - 1 + (slot & 1), ej = j - 1 + ((slot >> 1) & 1), ek = k - 1 + ((slot >> 2) & 1);
            if (ei < 0 || ej < 0 || ek < 0 || ei >= a.EI || ej >= a.EJ || ek >= a.EK) continue;
            const int e = ei + a.EI * (ej + a.EJ * ek);
            const DevMaterial& m = mat_of<MM>(a, e);
            const double inertialDensity = m.lam2mu / (c * c);
            double X[8][3];
            load_hex(a, a.xc, ei, ej, ek, X);
#pragma unroll
            for (int idx = 0; idx < 5; idx++) {
                const int t = kNodeToTetra[slot][idx];
                const int f = kNodeToSurface[slot][idx];
                /* effDensity = tetra->density (Force.c:109) */
                const double rho = TH ? a.dens[(size_t)t * E + e] : m.rho_eff;
                const double* p0 = X[kTetraToNode[t][0]];
                const double* p1 = X[kTetraToNode[t][1]];
                const double* p2 = X[kTetraToNode[t][2]];
                const double* p3 = X[kTetraToNode[t][3]];
                const double V = tet_volume(p0, p1, p2, p3);
                double ax, ay, az;
                face_cross(X[kTetraToNode[t][kFaceToNode[f][0]]], X[kTetraToNode[t][kFaceToNode[f][1]]],
                           X[kTetraToNode[t][kFaceToNode[f][2]]], ax, ay, az);
                const double sq = sqrt(ax * ax + ay * ay + az * az);
                const double area = sq / 2.0;
                const double ysq = 1.0 / sq;
                const double n0 = snh_div_by(ax, sq, ysq), n1 = snh_div_by(ay, sq, ysq), n2 = snh_div_by(az, sq, ysq);
                const double s00 = a.stress[((size_t)0 * NTET + t) * E + e];
                const double s11 = a.stress[((size_t)1 * NTET + t) * E + e];
                const double s22 = a.stress[((size_t)2 * NTET + t) * E + e];
                const double s01 = a.stress[((size_t)3 * NTET + t) * E + e];
                const double s02 = a.stress[((size_t)4 * NTET + t) * E + e];
                const double s12 = a.stress[((size_t)5 * NTET + t) * E + e];
                if (a.dtType == SNH_DT_DYNAMIC) {
                    im += 0.25 * inertialDensity * V;
                } else if (a.dtType == SNH_DT_CONSTANT) {
                    double area_sum = 0.0, tmp;
                    tmp = (n0 * area) * (n0 * area);
                    if (tmp > area_sum) area_sum = tmp;
                    tmp = (n1 * area) * (n1 * area);
                    if (tmp > area_sum) area_sum = tmp;
                    tmp = (n2 * area) * (n2 * area);
                    if (tmp > area_sum) area_sum = tmp;
                    im += m.lam2mu * area_sum / (9.0 * V) * (dt * dt);
                } else {
                    im += 0.25 * rho * V;
                }
                f0 += snh_div6((s00 * n0 + s01 * n1 + s02 * n2) * area);   /* (...)*area/6.0 */
                f1 += snh_div6((s01 * n0 + s11 * n1 + s12 * n2) * area);
                f2 += snh_div6((s02 * n0 + s12 * n1 + s22 * n2) * area);
                f1 -= a.gravity * rho * V / 8.0;
            }
        }
        vmag = node_finish(a, n, f0, f1, f2, im, own, dt);
    }
    double bm = block_max(vmag, sh);
    if (threadIdx.x == 0) a.part_vmax[blk] = bm;
}

template <bool MM, bool TH>
__global__ void __launch_bounds__(NODE_BLOCK) k_node(StepArgs a) {
    node_pass<MM, TH>(a, (int)blockIdx.x, (int)gridDim.x);
}

/* --------------------------------------------------------------------------
 * k_node_tile: the same node pass as k_node, organised so that each tet's
 * volume is computed once per tile instead of once per node of the tet, and
 * node coordinates come from LDS instead of 24 global loads per element.
 *
 * A workgroup owns NT_I x NT_J node columns (i, j) and a chunk of node
 * planes k0 <= k < k1, and walks up in k.  At element plane k it stages the
 * coordinates of node plane k+1 (planes k and k+1 sit in a two-slot ring)
 * and the ten tet volumes of every element of the plane that touches its
 * columns ((NT_I+1) x (NT_J+1) elements).  Node (i, j, k) then adds its
 * slots 4..7 (elements (i-1..i, j-1..j, k), dz = 1 in the reference's slot
 * order s = dx + 2 dy + 4 dz) to the slots 0..3 it took from plane k-1,
 * and is finished; node (i, j, k+1) takes its slots 0..3 from the same four
 * elements.  Each node therefore sums its 40 tet-face shares in exactly the
 * reference's order (Force.c:87-183), and every share is computed with the
 * same operations as in k_node: the two kernels give the same bits.  A
 * chunk that starts above plane 0 first takes slots 0..3 of its first node
 * plane from element plane k0-1.
 * ------------------------------------------------------------------------ */
#define NT_I 16
#define NT_J 8
#define NT_THREADS (NT_I * NT_J)
#define NT_CI (NT_I + 2)   /* node columns staged: one halo column each side */
#define NT_CJ (NT_J + 2)
#define NT_EI (NT_I + 1)   /* element columns touching the tile's nodes */
#define NT_EJ (NT_J + 1)

/* one tet face's share in Snac_Force (Force.c:108-173) for the node the face
   is opposite to: inertial mass, stress . n A / 6, gravity; acc = f0, f1,
   f2, inertial mass */
__device__ __forceinline__ void face_share(const StepArgs& a, const DevMaterial& m, double inertialDensity,
                                           double rho, double V, const double* c1, const double* c2,
                                           const double* c3, const double st[6], double dt, double acc[4]) {
    double ax, ay, az;
    face_cross(c1, c2, c3, ax, ay, az);
    const double sq = sqrt(ax * ax + ay * ay + az * az);
    const double area = sq / 2.0;
    const double ysq = 1.0 / sq;
    const double n0 = snh_div_by(ax, sq, ysq), n1 = snh_div_by(ay, sq, ysq), n2 = snh_div_by(az, sq, ysq);
    if (a.dtType == SNH_DT_DYNAMIC) {
        acc[3] += 0.25 * inertialDensity * V;
    } else if (a.dtType == SNH_DT_CONSTANT) {
        double area_sum = 0.0, tmp;
        tmp = (n0 * area) * (n0 * area);
        if (tmp > area_sum) area_sum = tmp;
        tmp = (n1 * area) * (n1 * area);
        if (tmp > area_sum) area_sum = tmp;
        tmp = (n2 * area) * (n2 * area);
        if (tmp > area_sum) area_sum = tmp;
        acc[3] += m.lam2mu * area_sum / (9.0 * V) * (dt * dt);
    } else {
        acc[3] += 0.25 * rho * V;
    }
    acc[0] += snh_div6((st[0] * n0 + st[3] * n1 + st[4] * n2) * area);
    acc[1] += snh_div6((st[3] * n0 + st[1] * n1 + st[5] * n2) * area);
    acc[2] += snh_div6((st[4] * n0 + st[5] * n1 + st[2] * n2) * area);
    acc[1] -= a.gravity * rho * V / 8.0;
}

/* the 40 tet-face shares of a node, per node-slot s and rank r in the
   reference's order: bits 0-3 the tet (NodeToTetra[s][r]); bits 4+3v..6+3v
   the hex offset (dx | dy << 1 | dz << 2) of vertex v of the tet face
   opposite the node (NodeToSurface[s][r], FaceToNode order) */
struct ShareTable {
    uint32_t v[8][5];
};
constexpr ShareTable make_share_table() {
    ShareTable tab{};
    for (int sl = 0; sl < 8; sl++)
        for (int r = 0; r < 5; r++) {
            const int t = kNodeToTetra[sl][r], f = kNodeToSurface[sl][r];
            uint32_t code = (uint32_t)t;
            for (int v = 0; v < 3; v++) {
                const int h = kTetraToNode[t][kFaceToNode[f][v]];
                code |= (uint32_t)(kHexOff[h][0] | (kHexOff[h][1] << 1) | (kHexOff[h][2] << 2)) << (4 + 3 * v);
            }
            tab.v[sl][r] = code;
        }
    return tab;
}
__constant__ ShareTable kShare = make_share_table();

template <bool MM, bool TH>
__global__ void __launch_bounds__(NT_THREADS) k_node_tile(StepArgs a) {
    __shared__ double sx[2][3][NT_CJ][NT_CI];     /* node planes k, k+1 (slot = plane & 1) */
    __shared__ double sv[NTET][NT_EJ][NT_EI];     /* tet volumes of element plane k */
    __shared__ double sh[NT_THREADS / 64];
    const int tid = threadIdx.x;
    const int ti = tid % NT_I, tj = tid / NT_I;
    const int nti = (a.I + NT_I - 1) / NT_I, ntj = (a.J + NT_J - 1) / NT_J;
    const int b = xcd_block();
    const int i0 = (b % nti) * NT_I, j0 = ((b / nti) % ntj) * NT_J;
    const int k0 = (b / (nti * ntj)) * a.node_kc;
    const int k1 = min(k0 + a.node_kc, a.K);
    const int i = i0 + ti, j = j0 + tj;
    const bool active = i < a.I && j < a.J;
    const double dt = a.scal->dt;
    const double c = a.scal->speedOfSound;
    const size_t E = (size_t)a.E, N = (size_t)a.N;

    auto load_plane = [&](int p) {
        double(*dst)[NT_CJ][NT_CI] = sx[p & 1];
        for (int q = tid; q < NT_CI * NT_CJ; q += NT_THREADS) {
            const int il = q % NT_CI, jl = q / NT_CI;
            const int gi = i0 - 1 + il, gj = j0 - 1 + jl;
            double x0 = 0.0, x1 = 0.0, x2 = 0.0;
            if (gi >= 0 && gi < a.I && gj >= 0 && gj < a.J) {
                const size_t n = (size_t)gi + (size_t)a.I * ((size_t)gj + (size_t)a.J * p);
                x0 = a.xc[n];
                x1 = a.xc[N + n];
                x2 = a.xc[2 * N + n];
            }
            dst[0][jl][il] = x0;
            dst[1][jl][il] = x1;
            dst[2][jl][il] = x2;
        }
    };
    auto volumes = [&](int ek) {
        const int s0 = ek & 1, s1 = s0 ^ 1;
        for (int q = tid; q < NT_EI * NT_EJ; q += NT_THREADS) {
            const int el = q % NT_EI, ej = q / NT_EI;
            const int gi = i0 - 1 + el, gj = j0 - 1 + ej;
            if (gi < 0 || gi >= a.EI || gj < 0 || gj >= a.EJ) continue;
            double X[8][3];
#pragma unroll
            for (int h = 0; h < 8; h++) {
                const int sl = kHexOff[h][2] ? s1 : s0;
#pragma unroll
                for (int d = 0; d < 3; d++) X[h][d] = sx[sl][d][ej + kHexOff[h][1]][el + kHexOff[h][0]];
            }
#pragma unroll
            for (int t = 0; t < NTET; t++)
                sv[t][ej][el] = tet_volume(X[kTetraToNode[t][0]], X[kTetraToNode[t][1]], X[kTetraToNode[t][2]],
                                           X[kTetraToNode[t][3]]);
        }
    };
    /* the five tet-face shares of node-slot `slot` of element e (local
       lowest node li, lj in the staged planes) into f0..im, in the
       reference's order; slot and rank are uniform across the workgroup, so
       the table entry is a scalar load and the LDS offsets scalar values */
    auto slot_shares = [&](int slot, int e, int li, int lj, int s0, int s1, const DevMaterial& m,
                           double inertialDensity, double& f0, double& f1, double& f2, double& im) {
        const double* b0 = &sx[s0][0][lj][li];   /* vertices with dz = 0 (plane ek) */
        const double* b1 = &sx[s1][0][lj][li];   /* dz = 1 (plane ek + 1) */
        const double* vb = &sv[0][lj][li];
        const double* sb = a.stress + e;
        /* the next share's stress is loaded while this one is computed */
        uint32_t code = kShare.v[slot][0];
        double st[6];
#pragma unroll
        for (int cc = 0; cc < 6; cc++) st[cc] = sb[((size_t)cc * NTET + (code & 15u)) * E];
#pragma unroll 1
        for (int idx = 0; idx < 5; idx++) {
            const int t = (int)(code & 15u);
            double stn[6];
            uint32_t code_n = 0;
            if (idx < 4) {
                code_n = kShare.v[slot][idx + 1];
#pragma unroll
                for (int cc = 0; cc < 6; cc++) stn[cc] = sb[((size_t)cc * NTET + (code_n & 15u)) * E];
            }
            const double rho = TH ? a.dens[(size_t)t * E + e] : m.rho_eff;
            const double V = vb[t * (NT_EJ * NT_EI)];
            double p[3][3];
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const uint32_t o = (code >> (4 + 3 * v)) & 7u;
                const double* pb = (o & 4u) ? b1 : b0;
                const int off = (int)((o >> 1) & 1u) * NT_CI + (int)(o & 1u);
#pragma unroll
                for (int d = 0; d < 3; d++) p[v][d] = pb[d * (NT_CJ * NT_CI) + off];
            }
            double acc[4] = {f0, f1, f2, im};
            face_share(a, m, inertialDensity, rho, V, p[0], p[1], p[2], st, dt, acc);
            f0 = acc[0];
            f1 = acc[1];
            f2 = acc[2];
            im = acc[3];
            if (idx < 4) {
#pragma unroll
                for (int cc = 0; cc < 6; cc++) st[cc] = stn[cc];
                code = code_n;
            }
        }
    };
    /* from the four elements of plane ek: when finish, slots 4..7 of node
       (i, j, ek) into fin; when start, slots 0..3 of node (i, j, ek+1) into
       nxt */
    auto gather = [&](int ek, bool finish, double fin[4], bool start, double nxt[4]) {
        const int s0 = ek & 1, s1 = s0 ^ 1;
#pragma unroll 1
        for (int sl = 0; sl < 4; sl++) {
            const int dx = sl & 1, dy = sl >> 1;
            const int ei = i - 1 + dx, ej = j - 1 + dy;
            if (ei < 0 || ej < 0 || ei >= a.EI || ej >= a.EJ) continue;
            const int e = ei + a.EI * (ej + a.EJ * ek);
            const int li = ti + dx, lj = tj + dy;   /* element's lowest node, in the staged planes */
            const DevMaterial& m = mat_of<MM>(a, e);
            const double inertialDensity = m.lam2mu / (c * c);
            if (finish) slot_shares(sl + 4, e, li, lj, s0, s1, m, inertialDensity, fin[0], fin[1], fin[2], fin[3]);
            if (start) slot_shares(sl, e, li, lj, s0, s1, m, inertialDensity, nxt[0], nxt[1], nxt[2], nxt[3]);
        }
    };

    double vmag = -INFINITY;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};   /* f0, f1, f2, inertial mass of node (i, j, k) */
    if (k0 >= 1) {
        load_plane(k0 - 1);
        load_plane(k0);
        __syncthreads();
        volumes(k0 - 1);
        __syncthreads();
        if (active) gather(k0 - 1, false, acc, true, acc);   /* slots 0..3 of node plane k0 only */
        __syncthreads();
    } else {
        load_plane(k0);
        __syncthreads();
    }
    for (int k = k0; k < k1; k++) {
        double nxt[4] = {0.0, 0.0, 0.0, 0.0};
        const bool up = k < a.K - 1;   /* element plane k exists */
        const bool start = k + 1 < k1;
        if (up) {
            load_plane(k + 1);
            __syncthreads();
            volumes(k);
            __syncthreads();
            if (active) gather(k, true, acc, start, nxt);
        }
        if (active) {
            const int sl = k & 1;
            const size_t n = (size_t)i + (size_t)a.I * ((size_t)j + (size_t)a.J * k);
            const double v = node_finish(a, (int)n, acc[0], acc[1], acc[2], acc[3], node_own(a, (int)n), dt);
            vmag = fmax(vmag, v);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] = nxt[q];
        __syncthreads();
    }
    const double bm = block_max(vmag, sh);
    if (tid == 0) a.part_vmax[blockIdx.x] = bm;
}

/* --------------------------------------------------------------------------
 * Topography diffusion of the top surface (topo_kappa > 0; UpdateNode.c:
 * 60-72,161-260,359-505,647-857, Cartesian branches).  The reference runs it
 * per node inside the second node pass, in node order, right after the
 * node's coordinate update: getMaxTopoGrad raises the running maximum of
 * |grad y| (never reset, Context.c:759-764), and once that reaches
 * topoGradCriterion, DiffTopo adds kappa dt lapl(y) to a top interior node
 * and AdjustEdgeTopo gives an edge node the height of its inner neighbour
 * (the mean of two at a corner).  A top node reads the final heights of its
 * -x and -z neighbours and the old coordinates of its +x and +z ones (slots
 * 3, 5 and 0, 2 of IJK6Topology_BuildNeighbours), so the final height of
 * node (i, k) depends on (i-1, k) and (i, k-1) only: the surface is swept
 * by anti-diagonals d = i + k.  Two nodes wide in x (resp. z) and more in
 * the other direction, the reference diffuses along z (resp. x) only
 * (diffTopo_alongZ/_alongX, getMaxTopoGrad_alongZ/_alongX and the last two
 * branches of AdjustEdgeTopo, UpdateNode.c:248-251,262-357,494-499,508-645,
 * 686-708,811-856): a node then depends on (i, k-1) (resp. (i-1, k)) alone,
 * so each of the two lines is a chain of its own, which k_topo_line runs
 * one thread per line; two nodes wide in both it does nothing.
 *
 * k_topo_prep (one thread per top node) computes everything the sweep does
 * not change -- the node's gradient for the running maximum, the
 * differences with its old +x/+z neighbours and the stencil's denominators
 * with their correctly rounded reciprocals -- into records stored in
 * diagonal order, and reduces the maximum and the first node (in node
 * order) whose gradient reaches the criterion per block.  k_topo (one
 * block) finishes the reductions and sweeps the diagonals with the heights
 * of diagonal d-1 in LDS, prefetching records several diagonals ahead
 * (written by k_topo_prep on all XCDs just before); a
 * division whose divisor the sweep does not change is done as
 * snh_div_by with the stored reciprocal (correctly rounded, as the
 * reference's).  Both run after k_node with its arguments: xc the old
 * coordinates, xn the updated ones, whose heights k_topo finalises.
 * ------------------------------------------------------------------------ */
#define TOPO_REC 11          /* doubles per top-node record            */
#define TOPO_PREP_BLOCK 256
#define TOPO_LINE_REC 8      /* doubles per record in the one-dimensional modes */

/* which of the reference's branches applies (on nodeGlobal3DCounts) */
enum { TOPO_SURFACE = 0, TOPO_ALONG_Z = 1, TOPO_ALONG_X = 2, TOPO_NONE = 3 };

struct TopoArgs {
    double* rec;             /* [TOPO_REC][I K] records, diagonal order;
                                one-dimensional modes: [lines][L][TOPO_LINE_REC] */
    double* pmax;            /* [nparts] block maxima of |grad y|       */
    int* pfirst;             /* [nparts] block minima of the first node */
    int nparts;
    int mode;                /* TOPO_SURFACE, _ALONG_Z or _ALONG_X      */
    double kappa, crit;
};

static int topo_mode(int I, int K) {
    if (I > 2 && K > 2) return TOPO_SURFACE;
    if (I == 2 && K > 2) return TOPO_ALONG_Z;
    if (K == 2 && I > 2) return TOPO_ALONG_X;
    return TOPO_NONE;
}

__device__ __forceinline__ int topo_ilo(int d, int K) { return d - (K - 1) > 0 ? d - (K - 1) : 0; }
/* first record of diagonal d: the top nodes with i + k < d on an I x K
   surface, sum over s < d of (min(s, I-1) + 1) - max(0, s - K + 1) */
__host__ __device__ __forceinline__ long long topo_off(int d, int I, int K) {
    const long long dd = d, II = I, KK = K;
    const long long a = dd <= II ? dd * (dd + 1) / 2 : II * (II + 1) / 2 + (dd - II) * II;
    const long long b = dd > KK ? (dd - KK) * (dd - KK + 1) / 2 : 0;
    return a - b;
}

__global__ void __launch_bounds__(TOPO_PREP_BLOCK) k_topo_prep(StepArgs a, TopoArgs t) {
    __shared__ double smax[TOPO_PREP_BLOCK];
    __shared__ int sfirst[TOPO_PREP_BLOCK];
    const int I = a.I, J = a.J, K = a.K, IJ = I * J, P = I * K;
    const int p = blockIdx.x * TOPO_PREP_BLOCK + threadIdx.x;   /* k I + i: node order */
    double g = 0.0;
    int first = P;
    if (p < P) {
        const int k = p / I, i = p - k * I;
        const int n = I * (J - 1) + i + IJ * k;
        const int d = i + k;
        const size_t N = (size_t)a.N, PP = (size_t)P;
        const size_t pos = (size_t)topo_off(d, I, K) + (size_t)(i - topo_ilo(d, K));
        double r[TOPO_REC];
#pragma unroll
        for (int q = 0; q < TOPO_REC; q++) r[q] = 0.0;
        r[0] = a.xn[N + n];
        const bool inx = i > 0 && i < I - 1, inz = k > 0 && k < K - 1;
        const bool alx = t.mode != TOPO_ALONG_Z, alz = t.mode != TOPO_ALONG_X;
        const bool interior = (!alx || inx) && (!alz || inz);
        if (interior) {
            const double y0 = r[0];
            double dydx1 = 0.0, dydz1 = 0.0;
            if (alx) {
                const double x0 = a.xn[n], xL = a.xn[n - 1];
                const double xR = a.xc[n + 1], yR = a.xc[N + n + 1];
                dydx1 = (yR - y0) / (xR - x0);
                r[1] = dydx1;
                r[2] = x0 - xL;
                r[3] = 1.0 / r[2];
                r[4] = 0.5 * (xR - xL);
                r[5] = 1.0 / r[4];
            }
            if (alz) {
                const double z0 = a.xn[2 * N + n], zB = a.xn[2 * N + n - IJ];
                const double yF = a.xc[N + n + IJ], zF = a.xc[2 * N + n + IJ];
                dydz1 = (yF - y0) / (zF - z0);
                r[6] = dydz1;
                r[7] = z0 - zB;
                r[8] = 1.0 / r[7];
                r[9] = 0.5 * (zF - zB);
                r[10] = 1.0 / r[9];
            }
            /* getMaxTopoGrad's dydx, dydz are the same quotients */
            g = (alx && alz) ? sqrt(dydx1 * dydx1 + dydz1 * dydz1)
                             : (alx ? sqrt(dydx1 * dydx1) : sqrt(dydz1 * dydz1));
            if (!(g < t.crit)) first = p;
        } else {
            if (alx && i == 0) r[1] = a.xc[N + n + 1];    /* old height of (1, k)  */
            if (alz && k == 0) r[6] = a.xc[N + n + IJ];   /* old height of (i, 1)  */
        }
        if (t.mode == TOPO_SURFACE) {
#pragma unroll
            for (int q = 0; q < TOPO_REC; q++) t.rec[q * PP + pos] = r[q];
        } else {
            /* one line per row (along x) or column (along z): the height and
               the direction's five fields, 64 bytes per node in line order */
            const size_t lpos = alx ? (size_t)p : (size_t)i * K + k;
            double* o = t.rec + TOPO_LINE_REC * lpos;
            const int f0 = alx ? 1 : 6;
            o[0] = r[0];
#pragma unroll
            for (int q = 0; q < 5; q++) o[1 + q] = r[f0 + q];
        }
    }
    smax[threadIdx.x] = g;
    sfirst[threadIdx.x] = first;
    __syncthreads();
    for (int w = TOPO_PREP_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const double o = smax[threadIdx.x + w];
            smax[threadIdx.x] = smax[threadIdx.x] > o ? smax[threadIdx.x] : o;
            const int of = sfirst[threadIdx.x + w];
            if (of < sfirst[threadIdx.x]) sfirst[threadIdx.x] = of;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        t.pmax[blockIdx.x] = smax[0];
        t.pfirst[blockIdx.x] = sfirst[0];
    }
}

/* The running maximum over k_topo_prep's block partials (one block):
   stores the new maximum and returns the first node, in node order, from
   which the diffusion runs -- 0 if an earlier step already reached the
   criterion, I K if this one does not.  rmax, rfirst: blockDim.x entries,
   a power of two. */
__device__ int topo_reduce(const StepArgs& a, const TopoArgs& t, double* rmax, int* rfirst) {
    const int tid = threadIdx.x, B = blockDim.x, P = a.I * a.K;
    const double prev = a.scal->topoGradMax;
    double m = 0.0;
    int f = P;
    for (int q = tid; q < t.nparts; q += B) {
        m = m > t.pmax[q] ? m : t.pmax[q];
        if (t.pfirst[q] < f) f = t.pfirst[q];
    }
    rmax[tid] = m;
    rfirst[tid] = f;
    __syncthreads();
    for (int w = B / 2; w > 0; w >>= 1) {
        if (tid < w) {
            rmax[tid] = rmax[tid] > rmax[tid + w] ? rmax[tid] : rmax[tid + w];
            if (rfirst[tid + w] < rfirst[tid]) rfirst[tid] = rfirst[tid + w];
        }
        __syncthreads();
    }
    const int pfirst = (prev < t.crit) ? rfirst[0] : 0;
    if (tid == 0) a.scal->topoGradMax = prev > rmax[0] ? prev : rmax[0];
    return pfirst;
}

/* BLOCK: the most threads the sweep may be launched with (a diagonal's
   nodes, at most min(I, K)); PF: diagonals prefetched ahead -- deeper with
   the smaller block, whose waves may hold more registers */
template <int BLOCK, int PF>
__global__ void __launch_bounds__(BLOCK) k_topo(StepArgs a, TopoArgs t) {
    constexpr int TOPO_PF = PF;
    __shared__ double yd[2][TOPO_MAX_I];
    __shared__ double rmax[BLOCK];
    __shared__ int rfirst[BLOCK];
    const int I = a.I, J = a.J, K = a.K, IJ = I * J, P = I * K;
    const int tid = threadIdx.x;
    const size_t N = (size_t)a.N, PP = (size_t)P;
    const double kd = t.kappa * a.scal->dt;

    /* the running maximum: reaches the criterion at node pfirst (node order) */
    const int pfirst = topo_reduce(a, t, rmax, rfirst);

    /* the sweep; thread tid holds node i = ilo(d) + tid of diagonal d.  The
       records of diagonal d + TOPO_PF are loaded while diagonal d is
       computed, always from a valid position (clamped), and the node's
       update is computed without branches (interior and edge formulas both
       evaluated, the right one selected), so that nothing waits on the
       loads before their use three barriers later. */
    const int D = I + K - 1;
    double ring[TOPO_PF][TOPO_REC];
#pragma unroll
    for (int r = 0; r < TOPO_PF; r++) {
        const int dd = r < D ? r : D - 1;
        const int cn = (dd < I - 1 ? dd : I - 1) - topo_ilo(dd, K) + 1;
        const size_t pos = (size_t)topo_off(dd, I, K) + (tid < cn ? tid : cn - 1);
#pragma unroll
        for (int q = 0; q < TOPO_REC; q++) ring[r][q] = t.rec[q * PP + pos];
    }
    for (int d0 = 0; d0 < D; d0 += TOPO_PF) {
#pragma unroll
        for (int r = 0; r < TOPO_PF; r++) {
            const int d = d0 + r;
            if (d >= D) break;
            const int ilo = topo_ilo(d, K);
            const int cnt = (d < I - 1 ? d : I - 1) - ilo + 1;
            const bool valid = tid < cnt;
            const int i = valid ? ilo + tid : ilo, k = d - i;
            const double* yprev = yd[(d + 1) & 1];   /* final heights of diagonal d-1, by i */
            const double yl = yprev[i > 0 ? i - 1 : 0];
            const double yb = yprev[i];
            const double* rc = ring[r];
            const double y0 = rc[0];
            /* DiffTopo + UpdateInteriorTopo */
            const double dydx2 = snh_div_by(y0 - yl, rc[2], rc[3]);
            const double d2ydx2 = snh_div_by(rc[1] - dydx2, rc[4], rc[5]);
            const double dydz2 = snh_div_by(y0 - yb, rc[7], rc[8]);
            const double d2ydz2 = snh_div_by(rc[6] - dydz2, rc[9], rc[10]);
            const double lapl_topo = d2ydx2 + d2ydz2;
            const double yi = y0 + kd * lapl_topo;
            /* AdjustEdgeTopo: xN1 along x, xN2 along z, the old height of a
               neighbour not yet updated (+x, +z), the final one otherwise */
            const bool h1 = (i == 0 || i == I - 1), h2 = (k == 0 || k == K - 1);
            const double y1 = (i == 0) ? rc[1] : yl;
            const double y2 = (k == 0) ? rc[6] : yb;
            const double ye = (h1 && !h2) ? y1 : ((!h1 && h2) ? y2 : 0.5 * (y1 + y2));
            const bool active = valid && k * I + i >= pfirst;
            const double y = active ? ((h1 || h2) ? ye : yi) : y0;
            if (valid) yd[d & 1][i] = y;
            if (active) a.xn[N + I * (J - 1) + i + IJ * k] = y;
            {
                const int dn = d + TOPO_PF < D ? d + TOPO_PF : D - 1;
                const int cn = (dn < I - 1 ? dn : I - 1) - topo_ilo(dn, K) + 1;
                const size_t pos = (size_t)topo_off(dn, I, K) + (tid < cn ? tid : cn - 1);
#pragma unroll
                for (int q = 0; q < TOPO_REC; q++) ring[r][q] = t.rec[q * PP + pos];
            }
            __syncthreads();
        }
    }
}

/* The one-dimensional modes: two lines of L nodes (rows k = 0, 1 along x
   when the mesh is two nodes wide in z, columns i = 0, 1 along z when it is
   two wide in x), each a chain in which node s depends on node s-1 of its
   own line only.  One thread per line keeps the previous final height in a
   register -- no LDS hand-over, no barrier -- and prefetches the records of
   PF nodes ahead; the node update is the sweep's, restricted to the line's
   direction.  One block of 64 threads (the reduction uses them all). */
template <int MODE, int PF>
__global__ void __launch_bounds__(64) k_topo_line(StepArgs a, TopoArgs t) {
    __shared__ double rmax[64];
    __shared__ int rfirst[64];
    const int I = a.I, J = a.J, K = a.K, IJ = I * J;
    const size_t N = (size_t)a.N;
    const double kd = t.kappa * a.scal->dt;
    const int pfirst = topo_reduce(a, t, rmax, rfirst);
    const int line = threadIdx.x;
    if (line >= 2) return;
    constexpr bool AX = MODE == TOPO_ALONG_X;
    const int L = AX ? I : K;
    /* the records of this line, TOPO_LINE_REC doubles per node: the height,
       the slope to the old + neighbour (the old height of that neighbour at
       the line's first node), the two denominators and their reciprocals */
    const double* lrec = (const double*)__builtin_assume_aligned(t.rec, 64) + (size_t)line * L * TOPO_LINE_REC;
    double* ytop = a.xn + N + (size_t)I * (J - 1) + (AX ? (size_t)IJ * line : (size_t)line);
    const size_t ystride = AX ? 1 : (size_t)IJ;
    double ring[PF][6];
#pragma unroll
    for (int r = 0; r < PF; r++) {
        const double* rp = lrec + (size_t)(r < L ? r : L - 1) * TOPO_LINE_REC;
#pragma unroll
        for (int q = 0; q < 6; q++) ring[r][q] = rp[q];
    }
    /* the body is unconditional (past the line's end it computes on the
       clamped last record and stores nothing): a branch around the loads
       makes every node wait for the loads just issued */
    double yl = 0.0;   /* final height of node s-1 */
    for (int s0 = 0; s0 < L; s0 += PF) {
#pragma unroll
        for (int r = 0; r < PF; r++) {
            const int s = s0 + r;
            const double* rc = ring[r];
            const double y0 = rc[0];
            const double dy2 = snh_div_by(y0 - yl, rc[2], rc[3]);
            const double lapl_topo = snh_div_by(rc[1] - dy2, rc[4], rc[5]);
            const double yi = y0 + kd * lapl_topo;
            /* AdjustEdgeTopo: the first node copies the old height of its
               + neighbour, the last the final height of its - neighbour */
            const double ye = (s == 0) ? rc[1] : yl;
            const int p = AX ? line * I + s : s * I + line;   /* node order */
            const bool active = s < L && p >= pfirst;
            const double y = active ? ((s == 0 || s == L - 1) ? ye : yi) : y0;
            if (active) ytop[(size_t)s * ystride] = y;
            yl = y;
            const double* rp = lrec + (size_t)(s + PF < L ? s + PF : L - 1) * TOPO_LINE_REC;
#pragma unroll
            for (int q = 0; q < 6; q++) ring[r][q] = rp[q];
        }
    }
}

/* |v| over constrained nodes without an update (TimeStepZero's Sync) */
__global__ void __launch_bounds__(NODE_BLOCK) k_vmax(StepArgs a) {
    __shared__ double sh[NODE_BLOCK / 64];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    double vmag = -INFINITY;
    if (n < a.N && a.bc[n]) {
        const size_t N = (size_t)a.N;
        double v0 = a.v[n], v1 = a.v[N + n], v2 = a.v[2 * N + n];
        vmag = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    }
    double bm = block_max(vmag, sh);
    if (threadIdx.x == 0) a.part_vmax[blockIdx.x] = bm;
}

/* --------------------------------------------------------------------------
 * State transfers: the device's rows [C][M] (one row per component, M nodes
 * or elements) and the host's records [M][C] (the reference's index order),
 * transposed on the device through a 64-record LDS tile so that both sides
 * move contiguous runs.  Stress rows are (component q, tet t) = q*10 + t;
 * its record column is t*6 + q.
 * ------------------------------------------------------------------------ */
enum { LAY_PLAIN = 0, LAY_STRESS = 1 };

template <int LAY>
__device__ __forceinline__ int row_of_col(int col) {
    return LAY == LAY_STRESS ? (col % 6) * NTET + col / 6 : col;
}

template <int C, int LAY>
__global__ void __launch_bounds__(256) k_rows_to_records(const double* __restrict__ in, double* __restrict__ out,
                                                         size_t M) {
    __shared__ double tile[C][65];
    const size_t m0 = (size_t)blockIdx.x * 64;
    for (int idx = threadIdx.x; idx < C * 64; idx += 256) {
        const int r = idx >> 6, j = idx & 63;
        if (m0 + j < M) tile[r][j] = in[(size_t)r * M + m0 + j];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < C * 64; idx += 256) {
        const int j = idx / C, col = idx - j * C;
        if (m0 + j < M) out[(m0 + j) * C + col] = tile[row_of_col<LAY>(col)][j];
    }
}

template <int C, int LAY>
__global__ void __launch_bounds__(256) k_records_to_rows(const double* __restrict__ in, double* __restrict__ out,
                                                         size_t M) {
    __shared__ double tile[C][65];
    const size_t m0 = (size_t)blockIdx.x * 64;
    for (int idx = threadIdx.x; idx < C * 64; idx += 256) {
        const int j = idx / C, col = idx - j * C;
        if (m0 + j < M) tile[row_of_col<LAY>(col)][j] = in[(m0 + j) * C + col];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < C * 64; idx += 256) {
        const int r = idx >> 6, j = idx & 63;
        if (m0 + j < M) out[(size_t)r * M + m0 + j] = tile[r][j];
    }
}

/* --------------------------------------------------------------------------
 * SnacTemperature: the energy entry points
 * ------------------------------------------------------------------------ */

/* unit normal and area of a triangle, as Tetrahedra_SurfaceNormal /
   Tetrahedra_SurfaceArea round them (Tetrahedra.c:70-113) */
__device__ __forceinline__ void face_normal_area(const double* p0, const double* p1, const double* p2, double n[3],
                                                 double& area) {
    double ax, ay, az;
    face_cross(p0, p1, p2, ax, ay, az);
    const double sq = sqrt(ax * ax + ay * ay + az * az);
    area = sq / 2.0;
    const double ysq = 1.0 / sq;
    n[0] = snh_div_by(ax, sq, ysq);
    n[1] = snh_div_by(ay, sq, ysq);
    n[2] = snh_div_by(az, sq, ysq);
}

/* tet surfaces whose mean normal is the wall normal of the heat-flux
   condition, per wall i = 0, i = EI-1, k = 0, k = EK-1
   (temperature/VariableConditions.c:253-308) */
static constexpr unsigned char kWallFaces[4][4][2] = {
    {{0, 2}, {1, 2}, {5, 0}, {6, 0}},
    {{2, 0}, {3, 0}, {7, 0}, {8, 0}},
    {{0, 1}, {2, 2}, {6, 1}, {7, 1}},
    {{1, 1}, {3, 2}, {5, 1}, {8, 1}}};

/* 0.25 (n_a + n_b + n_c + n_d) of wall W's four surfaces; the indices are
   compile-time constants, so X stays in registers */
template <int W>
__device__ __forceinline__ void wall_normal(const double X[8][3], double nb[3]) {
    double nn[4][3];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int t = kWallFaces[W][r][0], f = kWallFaces[W][r][1];
        double area;
        face_normal_area(X[kTetraToNode[t][kFaceToNode[f][0]]], X[kTetraToNode[t][kFaceToNode[f][1]]],
                         X[kTetraToNode[t][kFaceToNode[f][2]]], nn[r], area);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) nb[d] = 0.25 * (nn[0][d] + nn[1][d] + nn[2][d] + nn[3][d]);
}

/* Snac_EP_LoopElementsEnergy: Snac_HeatFlux (HeatFlux.c:63-155) for every
   tet of the element, then SnacTemperature_HeatFluxBC on the elements of the
   x and z walls (VariableConditions.c:227-353).  One thread per element. */
template <bool MM>
__global__ void __launch_bounds__(GEO_BLOCK) k_heat_elem(StepArgs a) {
    const int e = xcd_block() * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    int ei, ej, ek;
    elem_ijk(a, e, ei, ej, ek);
    const size_t E = (size_t)a.E;
    const DevMaterial& m = mat_of<MM>(a, e);
    double X[8][3], T[8];
    load_hex(a, a.xc, ei, ej, ek, X);
#pragma unroll
    for (int q = 0; q < 8; q++) T[q] = a.temp[hex_node_index(a, ei, ej, ek, q)];
    /* the last matching wall's normal: mean of four tet surface normals */
    int wall = -1;
    if (ei == 0) wall = 0;
    if (ei == a.EI - 1) wall = 1;
    if (ek == 0) wall = 2;
    if (ek == a.EK - 1) wall = 3;
    double nb[3] = {0.0, 0.0, 0.0};
    switch (wall) {
    case 0: wall_normal<0>(X, nb); break;
    case 1: wall_normal<1>(X, nb); break;
    case 2: wall_normal<2>(X, nb); break;
    case 3: wall_normal<3>(X, nb); break;
    default: break;
    }
    const double bulkm = m.lambda + snh_div3(2.0 * m.mu);
#pragma unroll
    for (int t = 0; t < NTET; t++) {
        const double* P[4] = {X[kTetraToNode[t][0]], X[kTetraToNode[t][1]], X[kTetraToNode[t][2]],
                              X[kTetraToNode[t][3]]};
        const double V = tet_volume(P[0], P[1], P[2], P[3]);
        const double yV = 1.0 / V;
        double n[4][3], A[4];
#pragma unroll
        for (int f = 0; f < 4; f++)
            face_normal_area(P[kFaceToNode[f][0]], P[kFaceToNode[f][1]], P[kFaceToNode[f][2]], n[f], A[f]);
        const double T0 = T[kTetraToNode[t][0]], T1 = T[kTetraToNode[t][1]];
        const double T2 = T[kTetraToNode[t][2]], T3 = T[kTetraToNode[t][3]];
        double S[3], hf[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            /* node k pairs with surface (k+1)%4, summed left to right */
            S[d] = T0 * n[1][d] * A[1] + T1 * n[2][d] * A[2] + T2 * n[3][d] * A[3] + T3 * n[0][d] * A[0];
            hf[d] = snh_div_by(snh_div3(1.0 * m.kcond * S[d]), V, yV);   /* 1.0f*k*(..)/3.0f/V */
        }
        if (wall >= 0) {
            double g[3];
#pragma unroll
            for (int d = 0; d < 3; d++) g[d] = snh_div_by(snh_div3(-1.0 * S[d]), V, yV);
            double Fn = 0.0;
            Fn += g[0] * nb[0];
            Fn += g[1] * nb[1];
            Fn += g[2] * nb[2];
#pragma unroll
            for (int d = 0; d < 3; d++) hf[d] -= Fn * nb[d];
        }
#pragma unroll
        for (int d = 0; d < 3; d++) a.hflux[((size_t)d * NTET + t) * E + e] = hf[d];
        if (a.tstress) {
            const int n0 = hex_node_index(a, ei, ej, ek, kTetraToNode[t][0]);
            const int n1 = hex_node_index(a, ei, ej, ek, kTetraToNode[t][1]);
            const int n2 = hex_node_index(a, ei, ej, ek, kTetraToNode[t][2]);
            const int n3 = hex_node_index(a, ei, ej, ek, kTetraToNode[t][3]);
            const double temp = 0.25 * (T0 + T1 + T2 + T3);
            const double temp0 = 0.25 * (a.temp0[n0] + a.temp0[n1] + a.temp0[n2] + a.temp0[n3]);
            a.tstress[(size_t)t * E + e] = -1.0 * m.alpha * bulkm * (temp - temp0);
        }
    }
}

/* Snac_EP_LoopNodesEnergy: Snac_Heat for every node (Heat.c:64-132, source
   term 0) with the node-centric gather of the force pass, then the
   temperature before the BCs (read by the tet averages of the next Stress
   call) and the BCs.  The lumped heat capacity uses the tet densities of the
   last Stress call; dt_thermal / dt_mech <= 1 raises thermal_err. */
template <bool MM>
__global__ void __launch_bounds__(NODE_BLOCK) k_heat_node(StepArgs a) {
    const int g = xcd_block() * blockDim.x + threadIdx.x;
    if (g >= a.N) return;
    int i, j, k;
    const int n = gather_node(a, g, i, j, k);
    const size_t E = (size_t)a.E;
    const double dt = a.scal->dt;
    const double minLen = a.scal->minLen;
    double energy = 0.0, lumpVolume = 0.0;
    bool bad = false;
#pragma unroll
    for (int slot = 0; slot < 8; slot++) {
        const int ei = i - 1 + (slot & 1), ej = j - 1 + ((slot >> 1) & 1), ek = k - 1 + ((slot >> 2) & 1);
        if (ei < 0 || ej < 0 || ek < 0 || ei >= a.EI || ej >= a.EJ || ek >= a.EK) continue;
        const int e = ei + a.EI * (ej + a.EJ * ek);
        const DevMaterial& m = mat_of<MM>(a, e);
        double X[8][3];
        load_hex(a, a.xc, ei, ej, ek, X);
#pragma unroll
        for (int idx = 0; idx < 5; idx++) {
            const int t = kNodeToTetra[slot][idx];
            const int f = kNodeToSurface[slot][idx];
            const double V = tet_volume(X[kTetraToNode[t][0]], X[kTetraToNode[t][1]], X[kTetraToNode[t][2]],
                                        X[kTetraToNode[t][3]]);
            double nf[3], area;
            face_normal_area(X[kTetraToNode[t][kFaceToNode[f][0]]], X[kTetraToNode[t][kFaceToNode[f][1]]],
                             X[kTetraToNode[t][kFaceToNode[f][2]]], nf, area);
            const double h0 = a.hflux[((size_t)0 * NTET + t) * E + e];
            const double h1 = a.hflux[((size_t)1 * NTET + t) * E + e];
            const double h2 = a.hflux[((size_t)2 * NTET + t) * E + e];
            energy += snh_div6((h0 * nf[0] + h1 * nf[1] + h2 * nf[2]) * area);   /* (..)*area/6.0f */
            /* the reference's (density == 0 ? formula : density) gives the
               stored value in both branches when beta = 0 */
            lumpVolume += m.cp * (a.dens[(size_t)t * E + e] * V) / 8.0;
        }
        const double ratio = sqrt(m.lam2mu / m.phsDensity) * minLen / (m.kcond / (m.cp * m.phsDensity));
        if (!(ratio > 1.0)) bad = true;
    }
    const double T = a.temp[n];
    const double Tn = T + (energy + 0.0) * -1.0 * dt / lumpVolume;
    a.temp0[n] = T;
    a.tpre[n] = Tn;
    const uint8_t code = a.tbc[n];
    a.temp[n] = code ? a.tbcv[code] : Tn;
    if (bad) a.scal->thermal_err = 1;
}

#include "snh_mixed.h"

/* --------------------------------------------------------------------------
 * k_geo: Snac_EP_LoopElementsMomentum's geometry and minLengthScale
 * ------------------------------------------------------------------------ */
__device__ __forceinline__ void geo_pass(StepArgs a, const int blk, const int nblk) {
    __shared__ double sh[GEO_BLOCK / 64];
    const int e = xcd_block(blk, nblk) * blockDim.x + threadIdx.x;
    double emin = INFINITY;
    if (e < a.E) {
        int ei, ej, ek;
        elem_ijk(a, e, ei, ej, ek);
        double X[8][3];
        load_hex(a, a.xc, ei, ej, ek, X);
        emin = F_1EP27;
#pragma unroll
        for (int t = 0; t < NTET; t++) {
            double Xt[4][3];
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int d = 0; d < 3; d++) Xt[q][d] = X[kTetraToNode[t][q]][d];
            /* min over faces of V / (sqrt(ssq)/2): sqrt, halving and the
               division are monotone, so the face with the largest sum of
               squares gives the minimum exactly */
            double ssq_max = 0.0, a0 = 0, b0 = 0, c0 = 0;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                double fa, fb, fc;
                face_cross(Xt[kFaceToNode[f][0]], Xt[kFaceToNode[f][1]], Xt[kFaceToNode[f][2]], fa, fb, fc);
                if (f == 0) {
                    a0 = fa;
                    b0 = fb;
                    c0 = fc;
                }
                double ssq = fa * fa + fb * fb + fc * fc;
                ssq_max = ssq > ssq_max ? ssq : ssq_max;
            }
            const double V = tet_volume_from_face0(Xt[0], Xt[3], a0, b0, c0);
            const double area = sqrt(ssq_max) / 2.0;
            if (area) {
                double ls = fabs(V / area);
                if (emin > ls) emin = ls;
            }
        }
    }
    double bm = block_min(emin, sh);
    if (threadIdx.x == 0) {
        a.part_min[blk] = bm;
        if (blk == 0) a.scal->dt_rot = a.scal->dt;
    }
}

__global__ void __launch_bounds__(GEO_BLOCK) k_geo(StepArgs a) { geo_pass(a, (int)blockIdx.x, (int)gridDim.x); }

/* --------------------------------------------------------------------------
 * k_fin: _Snac_Context_Sync
 * ------------------------------------------------------------------------ */
__device__ __forceinline__ void fin_pass(StepArgs a, const int nmin, const int nvmax) {
    __shared__ double sh[FIN_BLOCK / 64];
    double mn = INFINITY, mx = -INFINITY;
    /* eight partials in flight per thread: indices past the end read the
       last partial again, which leaves a minimum or maximum unchanged (and
       neither depends on the order) */
    const int nb = (int)blockDim.x;
    for (int q0 = threadIdx.x; q0 < nmin; q0 += 8 * nb) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = a.part_min[min(q0 + u * nb, nmin - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) mn = fmin(mn, v[u]);
    }
    for (int q0 = threadIdx.x; q0 < nvmax; q0 += 8 * nb) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = a.part_vmax[min(q0 + u * nb, nvmax - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) mx = fmax(mx, v[u]);
    }
    mn = block_min(mn, sh);
    __syncthreads();
    mx = block_max(mx, sh);
    if (threadIdx.x == 0) {
        double vmax = mx > 0.0 ? mx : 0.0;   /* vmax starts at 0 (Context.c:1218) */
        if (vmax < F_3EM11) vmax = F_3EM11;
        DevScalars* s = a.scal;
        s->vmax = vmax;
        s->minLen = mn;
        if (a.dtType == SNH_DT_DYNAMIC) {
            s->dt = mn * F_045 * a.strain_inert / vmax;
            s->speedOfSound = mn * 0.5 / s->dt;
        } else if (a.dtType == SNH_DT_WAVE) {   /* Context.c:1255-1259 */
            s->speedOfSound = a.c_wave;
            s->dt = 0.5 * mn / s->speedOfSound;
        }
        if (!isfinite(s->dt)) s->nonfinite = 1;
    }
}

__global__ void __launch_bounds__(FIN_BLOCK) k_fin(StepArgs a, int nmin, int nvmax) { fin_pass(a, nmin, nvmax); }

/* --------------------------------------------------------------------------
 * Batched launches of an ensemble: one grid over the blocks of several
 * contexts.  Block b of the grid belongs to member member[b], as that
 * member's block b - first[member]; first[] is a multiple of 8, so every
 * member's blocks keep the XCD numbering they have in a launch of their own
 * (xcd_block), and the blocks past a member's count (padding to the next
 * multiple of 8) return at once.  Each member's pass is the one of its own
 * launch, on its own StepArgs, so the results are the same bits.
 * ------------------------------------------------------------------------ */
/* A member's StepArgs relative to the launch's first member: each pointer
   as a byte offset from that member's (null pointers included: offsets are
   integer differences), so that the kernel forms it from a kernel argument
   and the compiler keeps global-memory loads (a pointer read from memory
   would be a generic one, with flat loads); the sizes and constants as
   they are. */
struct BatchMember {
    int64_t xc;
    int64_t xp;
    int64_t xn;
    int64_t v;
    int64_t v_out;
    int64_t stress;
    int64_t stress_out;
    int64_t ps;
    int64_t ps_out;
    int64_t visc;
    int64_t mat;
    int64_t refvisc;
    int64_t cohoff;
    int64_t bc;
    int64_t tape_f;
    int64_t tape_m;
    int64_t tape_code;
    int64_t damp_sched;
    int64_t diag;
    int64_t force_out;
    int64_t temp;
    int64_t temp0;
    int64_t tpre;
    int64_t hflux;
    int64_t tstress;
    int64_t dens;
    int64_t tbc;
    int64_t tbcv;
    int64_t mats;
    int64_t bcs;
    int64_t part_min;
    int64_t part_vmax;
    int64_t scal;
    int32_t I, J, K, EI, EJ, EK, N, E, dtType, node_band, node_kc;
    FastDiv dI, dJ, dEI, dEJ, dKB, dPer;   /* StepArgs' divisors */
    double gravity, demf, strain_inert, c_wave;
};

struct BatchGrid {
    StepArgs base;                          /* the first member, this step's parity */
    const BatchMember* __restrict__ mem;    /* [members], same parity               */
    const int* __restrict__ first;          /* [members] first block                */
    const int* __restrict__ count;          /* [members] blocks of the member       */
    const uint16_t* __restrict__ member;    /* [grid] member of each block          */
};

__device__ __forceinline__ StepArgs member_args(const StepArgs& base, const BatchMember* __restrict__ mem,
                                                const int m) {
    StepArgs a = base;
    const BatchMember& d = mem[m];
#define SNH_REBASE(f) a.f = (decltype(a.f))((const char*)base.f + d.f)
    SNH_REBASE(xc);
    SNH_REBASE(xp);
    SNH_REBASE(xn);
    SNH_REBASE(v);
    SNH_REBASE(v_out);
    SNH_REBASE(stress);
    SNH_REBASE(stress_out);
    SNH_REBASE(ps);
    SNH_REBASE(ps_out);
    SNH_REBASE(visc);
    SNH_REBASE(mat);
    SNH_REBASE(refvisc);
    SNH_REBASE(cohoff);
    SNH_REBASE(bc);
    SNH_REBASE(tape_f);
    SNH_REBASE(tape_m);
    SNH_REBASE(tape_code);
    SNH_REBASE(damp_sched);
    SNH_REBASE(diag);
    SNH_REBASE(force_out);
    SNH_REBASE(temp);
    SNH_REBASE(temp0);
    SNH_REBASE(tpre);
    SNH_REBASE(hflux);
    SNH_REBASE(tstress);
    SNH_REBASE(dens);
    SNH_REBASE(tbc);
    SNH_REBASE(tbcv);
    SNH_REBASE(mats);
    SNH_REBASE(bcs);
    SNH_REBASE(part_min);
    SNH_REBASE(part_vmax);
    SNH_REBASE(scal);
#undef SNH_REBASE
    a.I = d.I;
    a.J = d.J;
    a.K = d.K;
    a.EI = d.EI;
    a.EJ = d.EJ;
    a.EK = d.EK;
    a.N = d.N;
    a.E = d.E;
    a.dtType = d.dtType;
    a.node_band = d.node_band;
    a.node_kc = d.node_kc;
    a.gravity = d.gravity;
    a.demf = d.demf;
    a.strain_inert = d.strain_inert;
    a.c_wave = d.c_wave;
    a.dI = d.dI;
    a.dJ = d.dJ;
    a.dEI = d.dEI;
    a.dEJ = d.dEJ;
    a.dKB = d.dKB;
    a.dPer = d.dPer;
    return a;
}

/* the member of this block (uniform: an SGPR) and its block index */
__device__ __forceinline__ int batch_member(const BatchGrid& g, int& b) {
    const int m = __builtin_amdgcn_readfirstlane((int)g.member[blockIdx.x]);
    b = (int)blockIdx.x - g.first[m];
    return m;
}

template <int RSET, bool TW, bool MM>
__global__ void __launch_bounds__(TW ? 640 : 64) k_elem_batch(BatchGrid g) {
    int b;
    const int m = batch_member(g, b);
    if (b >= g.count[m]) return;
    elem_pass<RSET, TW, MM>(member_args(g.base, g.mem, m), b, g.count[m]);
}

template <int RSET, bool TW>
__global__ void __launch_bounds__(TW ? 640 : 64) k_elem_mx_batch(BatchGrid g) {
    int b;
    const int m = batch_member(g, b);
    if (b >= g.count[m]) return;
    elem_pass_mx<RSET, TW>(member_args(g.base, g.mem, m), b, g.count[m]);
}

template <bool MM>
__global__ void __launch_bounds__(NODE_BLOCK) k_node_batch(BatchGrid g) {
    int b;
    const int m = batch_member(g, b);
    if (b >= g.count[m]) return;
    node_pass<MM, false>(member_args(g.base, g.mem, m), b, g.count[m]);
}

__global__ void __launch_bounds__(NODE_BLOCK) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_node_mx_batch(BatchGrid g) {
    int b;
    const int m = batch_member(g, b);
    if (b >= g.count[m]) return;
    node_pass_mx(member_args(g.base, g.mem, m), b, g.count[m]);
}

__global__ void __launch_bounds__(GEO_BLOCK) k_geo_batch(BatchGrid g) {
    int b;
    const int m = batch_member(g, b);
    if (b >= g.count[m]) return;
    geo_pass(member_args(g.base, g.mem, m), b, g.count[m]);
}

/* block m: member m's Sync, over its geo and node block partials */
__global__ void __launch_bounds__(FIN_BLOCK) k_fin_batch(StepArgs base, const BatchMember* __restrict__ mem,
                                                         const int* __restrict__ nmin,
                                                         const int* __restrict__ nvmax) {
    const int m = (int)blockIdx.x;
    fin_pass(member_args(base, mem, m), nmin[m], nvmax[m]);
}

/* minLengthScale after a remesh: the running minimum of
   _SnacRemesher_UpdateElements (remesher_BI/RemeshElements.c:133-151) starts
   from the current value; dt is not recomputed */
__global__ void __launch_bounds__(FIN_BLOCK) k_minlen_merge(StepArgs a, int nmin) {
    __shared__ double sh[FIN_BLOCK / 64];
    double mn = INFINITY;
    for (int q = threadIdx.x; q < nmin; q += blockDim.x) mn = fmin(mn, a.part_min[q]);
    mn = block_min(mn, sh);
    if (threadIdx.x == 0 && mn < a.scal->minLen) a.scal->minLen = mn;
}

#include "snh_adjoint.h"

/* prescribed dt and speed of sound of one step (snh_set_schedule) */
__global__ void k_set_scalars(DevScalars* __restrict__ s, const double* __restrict__ dt,
                              const double* __restrict__ cs) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->dt = *dt;
        s->speedOfSound = *cs;
    }
}

/* ==========================================================================
 * host side
 * ======================================================================== */
static thread_local char g_err[512];

static int set_err(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

/* a roctx range over one entry point (rocprofv3 --marker-trace shows them;
   without a tool attached they cost a function call) */
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
};

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return set_err(SNH_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                           __LINE__);                                                             \
    } while (0)

struct snh_context {
    snh_config cfg;
    std::vector<snh_material> mats;
    std::vector<snh_velocity_bc> bcs;
    int I, J, K, EI, EJ, EK, N, E;
    hipStream_t stream;
    bool own_stream;
    /* device buffers */
    double* x[2];
    int cur;        /* index of the current coordinate buffer */
    double* v;
    double* stress;
    double* ps;
    double* visc;
    uint8_t* mat;
    double* refvisc;
    double* cohoff;
    uint8_t* bc;
    double* part_min;
    double* part_vmax;
    DevScalars* scal;
    DevMaterial* dmat;   /* material table [SNH_MAX_MATERIALS] */
    DevBC* dbc;          /* velocity-condition table [256], entry 0 unused */
    double* tmp;    /* download staging [6][10][E] */
    double* stage;  /* host-order records of one field, max(60 E, 3 N) doubles */
    /* topography diffusion (topo_kappa > 0): records and partial reductions
       of k_topo_prep / k_topo */
    double* topo_rec;
    double* topo_pmax;
    int* topo_pfirst;
    int topo_parts;
    int topo_threads;   /* sweep block: a power of two >= min(I, K) */
    int n_geo_blocks, n_node_blocks;
    unsigned rset;  /* rheologies present */
    bool zero_done;
    bool rotation_pending;
    bool needs_plastic;
    bool needs_visc;
    /* Vp = sqrt((lambda + 2 mu) / phsDensity) at its maximum over the
       elements' materials: the wave dt type's speed of sound
       (UpdateElement.c:62-67, a running maximum that never changes after
       the first element pass) */
    double c_wave;
    /* a material with beta != 0: tet densities depend on the pressure
       (k_density after every element pass) */
    bool beta_dens;
    uint32_t timeStep;
    int elem_variant;   /* SNH_ELEM_* */
    int node_band;      /* StepArgs::node_band */
    int node_variant;   /* SNH_NODE_* */
    int node_kc;        /* StepArgs::node_kc */
    int n_tile_blocks;  /* workgroups of k_node_tile */
    int n_vmax_parts;   /* block maxima of |v| written by the last node pass */
    struct AdjointCtx* adj;   /* recorded adjoint run, or null */
    bool diag_on;             /* steps record dump diagnostics */
    double* diag_buf;         /* [5][E] */
    double* force_buf;        /* [3][N] */
    double* diag_mx;          /* mixed precision: the FP64 diagnostic pass's
                                 stress [60E] and plastic strain [10E] */
    /* SnacTemperature (snh_set_thermal) */
    bool thermal;
    bool thermal_stress;
    double* temp;             /* [N] */
    double* temp0;            /* [N] */
    double* tpre;             /* [N] */
    double* hflux;            /* [3][10][E] */
    double* tstress;          /* [10][E] or null */
    double* dens;             /* [10][E] */
    uint8_t* tbc;             /* [N] */
    double* tbcv;             /* [256] */
    /* prescribed dt / speed of sound / damping branch for steps
       [sched_base, sched_base + sched_n) of timeStep */
    double* sched_dt;
    double* sched_c;
    uint8_t* sched_damp;      /* [sched_n][N] or null */
    uint32_t sched_base;
    int32_t sched_n;
    uint8_t* rec_code;        /* [N] damping branches of a recorded step (snh_record_schedule) */
    bool recording;           /* the node pass writes rec_code */
    /* kernel timing */
    bool profiling;
    std::vector<hipEvent_t> ev_pool;
    struct Rec { int kind; hipEvent_t a, b; };
    std::vector<Rec> recs;
    size_t ev_used;
};

/* Host transfers and fills of a context's device buffers, ordered on the
   context's stream: that stream is created non-blocking (or handed in by
   the caller), so it does not wait for the null stream, and a plain
   hipMemcpy / hipMemset there could overlap the stream's kernels.  A copy
   runs after the stream's earlier work and is complete on return (the host
   buffer may be reused, downloaded data may be read); a fill is ordered
   before the stream's later work. */
static hipError_t ctx_copy(snh_context* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

static hipError_t ctx_fill(snh_context* c, void* dst, int value, size_t bytes) {
    return hipMemsetAsync(dst, value, bytes, c->stream);
}

enum { KK_ELEM = 0, KK_NODE = 1, KK_GEO = 2, KK_FIN = 3, KK_HEAT_ELEM = 4, KK_HEAT_NODE = 5, KK_KINDS = 6 };

static hipEvent_t take_event(snh_context* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

/* brackets one launch with events when profiling is on */
struct KTimer {
    snh_context* c;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(snh_context* c_, int k) : c(c_), kind(k) {
        if (c->profiling) {
            a = take_event(c);
            b = take_event(c);
            if (a && b) (void)hipEventRecord(a, c->stream);
        }
    }
    ~KTimer() {
        if (c->profiling && a && b) {
            (void)hipEventRecord(b, c->stream);
            c->recs.push_back({kind, a, b});
        }
    }
};

extern "C" int snh_abi_version(void) { return SNH_ABI_VERSION; }
extern "C" const char* snh_last_error(void) { return g_err; }

static void fill_dev_material(const snh_material& s, DevMaterial& d) {
    memset(&d, 0, sizeof d);
    d.rheology = s.rheology;
    d.lambda = s.lambda;
    d.mu = s.mu;
    /* Stress.c:169-170 with avgTemp = 0 and beta = 0 */
    d.rho_eff = s.phsDensity * (1.0 - s.alpha * (0.0 - s.reftemp));
    d.lam2mu = s.lambda + 2.0 * s.mu;
    d.nseg = s.nsegments;
    for (unsigned i = 0; i <= s.nsegments && i < 9; i++) {
        d.plstrain[i] = s.plstrain[i];
        d.fa[i] = s.frictionAngle[i];
        d.da[i] = s.dilationAngle[i];
        d.coh[i] = s.cohesion[i];
    }
    for (unsigned i = 0; i < s.nsegments && i < 8; i++) {
        const double pl1 = s.plstrain[i], pl2 = s.plstrain[i + 1];
        d.tgf[i] = (s.frictionAngle[i + 1] - s.frictionAngle[i]) / (pl2 - pl1);
        d.tgd[i] = (s.dilationAngle[i + 1] - s.dilationAngle[i]) / (pl2 - pl1);
        d.tgc[i] = (s.cohesion[i + 1] - s.cohesion[i]) / (pl2 - pl1);
    }
    d.ten_off = s.ten_off;
    const double degrad = M_PI / 180.0;
    d.sphi = sin(s.frictionAngle[0] * degrad);
    d.spsi = sin(s.dilationAngle[0] * degrad);
    d.tanphi = tan(s.frictionAngle[0] * degrad);
    d.anphi0 = (1.0 + d.sphi) / (1.0 - d.sphi);
    d.sqanphi0 = sqrt(d.anphi0);
    d.vis_min = s.vis_min;
    d.vis_max = s.vis_max;
    d.refvisc = s.refvisc;
    d.refsrate = s.refsrate;
    d.reftemp = s.reftemp;
    d.activationE = s.activationE;
    d.srexponent = s.srexponent;
    d.linear_visc = (s.srexponent == 1.0 && s.activationE == 0.0);
    d.phsDensity = s.phsDensity;
    d.alpha = s.alpha;
    d.beta = s.beta;
    d.kcond = s.thermal_conduct;
    d.cp = s.heatCapacity;
}

static bool on_wall(const snh_context* c, int n, int wall) {
    int i = n % c->I, j = (n / c->I) % c->J, k = n / (c->I * c->J);
    switch (wall) {
    case SNH_WALL_FRONT: return c->K > 1 && k == 0;
    case SNH_WALL_BACK: return c->K > 1 && k == c->K - 1;
    case SNH_WALL_LEFT: return c->I > 1 && i == 0;
    case SNH_WALL_RIGHT: return c->I > 1 && i == c->I - 1;
    case SNH_WALL_TOP: return c->J > 1 && j == c->J - 1;
    case SNH_WALL_BOTTOM: return c->J > 1 && j == 0;
    case SNH_WALL_ALL: return true;
    default: return false;
    }
}

/* the multiply-high divisors of a StepArgs' index arithmetic */
static void set_divisors(StepArgs& a) {
    a.dI = fastdiv_make((uint32_t)a.I);
    a.dJ = fastdiv_make((uint32_t)a.J);
    a.dEI = fastdiv_make((uint32_t)(a.EI > 0 ? a.EI : 1));
    a.dEJ = fastdiv_make((uint32_t)(a.EJ > 0 ? a.EJ : 1));
    a.dKB = fastdiv_make((uint32_t)(a.node_band > 0 ? a.node_band : 1));
    a.dPer = fastdiv_make((uint32_t)(a.node_band > 0 ? a.J * a.node_band : 1));
}

static StepArgs make_args(snh_context* c) {
    StepArgs a;
    a.xc = c->x[c->cur];
    a.xp = c->x[c->cur ^ 1];
    a.xn = c->x[c->cur ^ 1];
    a.v = c->v;
    a.v_out = c->v;
    a.stress = c->stress;
    a.stress_out = c->stress;
    a.ps = c->ps;
    a.ps_out = c->ps;
    a.tape_f = nullptr;
    a.tape_m = nullptr;
    a.tape_code = nullptr;
    a.damp_sched = nullptr;
    a.diag = c->diag_on ? c->diag_buf : nullptr;
    a.force_out = c->diag_on ? c->force_buf : nullptr;
    a.visc = c->visc;
    a.mat = c->mat;
    a.refvisc = c->refvisc;
    a.cohoff = c->cohoff;
    a.bc = c->bc;
    a.part_min = c->part_min;
    a.part_vmax = c->part_vmax;
    a.scal = c->scal;
    a.mats = c->dmat;
    a.bcs = c->dbc;
    a.temp = c->thermal ? c->temp : nullptr;
    a.temp0 = c->thermal ? c->temp0 : nullptr;
    a.tpre = c->thermal ? c->tpre : nullptr;
    a.hflux = c->thermal ? c->hflux : nullptr;
    a.tstress = c->thermal && c->thermal_stress ? c->tstress : nullptr;
    a.dens = (c->thermal || c->beta_dens) ? c->dens : nullptr;
    a.tbc = c->thermal ? c->tbc : nullptr;
    a.tbcv = c->thermal ? c->tbcv : nullptr;
    a.I = c->I; a.J = c->J; a.K = c->K;
    a.EI = c->EI; a.EJ = c->EJ; a.EK = c->EK;
    a.N = c->N; a.E = c->E;
    a.dtType = c->cfg.dtType;
    a.node_band = c->node_band;
    a.node_kc = c->node_kc;
    a.gravity = c->cfg.gravity;
    a.demf = c->cfg.demf;
    a.strain_inert = c->cfg.strain_inert;
    a.c_wave = c->c_wave;
    set_divisors(a);
    return a;
}

/* Snac_UpdateElementMomentum's running maximum of Vp (UpdateElement.c:62-67) */
static double wave_speed_max(double cur, const snh_material& m) {
    const double Vp = sqrt((m.lambda + 2.0 * m.mu) / m.phsDensity);
    return Vp > cur ? Vp : cur;
}

static void free_all(snh_context* c) {
    void* ptrs[] = {c->x[0], c->x[1], c->v, c->stress, c->ps, c->visc, c->mat, c->refvisc, c->cohoff,
                    c->bc, c->part_min, c->part_vmax, c->scal, c->tmp, c->sched_dt, c->sched_c, c->sched_damp,
                    c->dmat, c->dbc, c->diag_buf, c->force_buf, c->diag_mx, c->temp, c->temp0, c->tpre, c->hflux,
                    c->tstress, c->dens, c->tbc, c->tbcv, c->stage, c->topo_rec, c->topo_pmax,
                    c->topo_pfirst, c->rec_code};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

extern "C" int snh_context_create(const snh_config* cfg, snh_context** out) {
    if (!cfg || !out) return set_err(SNH_EINVAL, "snh_context_create: null argument");
    *out = nullptr;
    if (cfg->meshSizeI < 2 || cfg->meshSizeJ < 2 || cfg->meshSizeK < 2)
        return set_err(SNH_EINVAL, "mesh needs at least 2 nodes per axis");
    if (cfg->nmaterials < 1 || cfg->nmaterials > SNH_MAX_MATERIALS || !cfg->materials)
        return set_err(SNH_EINVAL, "need 1..%d materials", SNH_MAX_MATERIALS);
    if (cfg->topo_kappa > 0) {
        /* the surface sweep keeps two rows of heights in LDS; the
           one-dimensional modes keep none */
        if (topo_mode((int)cfg->meshSizeI, (int)cfg->meshSizeK) == TOPO_SURFACE) {
            if (cfg->meshSizeI > TOPO_MAX_I)
                return set_err(SNH_EUNSUPPORTED, "topo_kappa > 0 with more than %d nodes along x", TOPO_MAX_I);
            if (cfg->meshSizeI > 1024 && cfg->meshSizeK > 1024)
                return set_err(SNH_EUNSUPPORTED, "topo_kappa > 0 with more than 1024 nodes along both x and z");
        }
        if (cfg->precision != SNH_PRECISION_FP64)
            return set_err(SNH_EUNSUPPORTED, "topo_kappa > 0 runs on the FP64 path");
    }
    if (cfg->dtType != SNH_DT_DYNAMIC && cfg->dtType != SNH_DT_CONSTANT && cfg->dtType != SNH_DT_WAVE)
        return set_err(SNH_EINVAL, "bad dtType");
    if (cfg->precision != SNH_PRECISION_FP64 && cfg->precision != SNH_PRECISION_MIXED)
        return set_err(SNH_EINVAL, "bad precision");
    if (cfg->precision == SNH_PRECISION_MIXED) {
        if (cfg->dtType != SNH_DT_DYNAMIC)
            return set_err(SNH_EUNSUPPORTED, "mixed precision runs the dynamic dt type only");
        for (unsigned q = 0; q < cfg->nmaterials; q++)
            if ((cfg->materials[q].rheology & SNH_RHEOLOGY_VISCOPLASTIC) &&
                !(cfg->materials[q].rheology & SNH_RHEOLOGY_PLASTIC))
                return set_err(SNH_EUNSUPPORTED, "mixed precision has no visco-plastic law yet");
    }
    if (cfg->nbcs && !cfg->bcs) return set_err(SNH_EINVAL, "nbcs > 0 with null bcs");
    double vol = (double)cfg->meshSizeI * cfg->meshSizeJ * cfg->meshSizeK;
    if (vol > 2.0e9) return set_err(SNH_EINVAL, "mesh too large for 32-bit node indices");
    for (unsigned q = 0; q < cfg->nmaterials; q++) {
        const snh_material& m = cfg->materials[q];
        if (m.beta != 0.0 && cfg->precision != SNH_PRECISION_FP64)
            return set_err(SNH_EUNSUPPORTED, "material %u: beta != 0 (pressure-dependent density) runs on the FP64 path",
                           q);
        if (m.nsegments > SNH_MAX_SEGMENTS) return set_err(SNH_EINVAL, "material %u: too many segments", q);
        for (unsigned s = 0; s <= m.nsegments; s++)
            if ((m.rheology & (SNH_RHEOLOGY_PLASTIC | SNH_RHEOLOGY_VISCOPLASTIC)) && m.frictionAngle[s] < 0)
                return set_err(SNH_EINVAL, "material %u: negative friction angle", q);
    }

    snh_context* c = new (std::nothrow) snh_context();
    if (!c) return set_err(SNH_ENOMEM, "out of host memory");
    c->cfg = *cfg;
    c->mats.assign(cfg->materials, cfg->materials + cfg->nmaterials);
    c->c_wave = wave_speed_max(0.0, c->mats[0]);   /* every element has material 0 until told otherwise */
    c->beta_dens = false;
    for (const snh_material& m : c->mats) c->beta_dens = c->beta_dens || m.beta != 0.0;
    if (cfg->nbcs) c->bcs.assign(cfg->bcs, cfg->bcs + cfg->nbcs);
    c->cfg.materials = c->mats.data();
    c->cfg.bcs = c->bcs.empty() ? nullptr : c->bcs.data();
    c->I = cfg->meshSizeI; c->J = cfg->meshSizeJ; c->K = cfg->meshSizeK;
    c->EI = c->I - 1; c->EJ = c->J - 1; c->EK = c->K - 1;
    c->N = c->I * c->J * c->K;
    c->E = c->EI * c->EJ * c->EK;
    c->rset = 0;
    for (auto& m : c->mats) {
        if (m.rheology & SNH_RHEOLOGY_PLASTIC) c->rset |= RS_PLASTIC;
        else if (m.rheology & SNH_RHEOLOGY_VISCOPLASTIC) c->rset |= RS_VISCOPLASTIC;
        else if (m.rheology & SNH_RHEOLOGY_MAXWELL) c->rset |= RS_MAXWELL;
        else if (m.rheology & SNH_RHEOLOGY_ELASTIC) c->rset |= RS_ELASTIC;
        /* the law of a material is its highest rheology bit, matching one
           rheology plugin per run */
    }
    for (auto& m : c->mats) {
        if (m.rheology & SNH_RHEOLOGY_PLASTIC) m.rheology = SNH_RHEOLOGY_PLASTIC;
        else if (m.rheology & SNH_RHEOLOGY_VISCOPLASTIC) m.rheology = SNH_RHEOLOGY_VISCOPLASTIC;
        else if (m.rheology & SNH_RHEOLOGY_MAXWELL) m.rheology = SNH_RHEOLOGY_MAXWELL;
    }
    c->needs_plastic = (c->rset & (RS_PLASTIC | RS_VISCOPLASTIC)) != 0;
    c->needs_visc = false;
    for (auto& m : c->mats)
        if ((m.rheology & (SNH_RHEOLOGY_MAXWELL | SNH_RHEOLOGY_VISCOPLASTIC)) &&
            !(m.srexponent == 1.0 && m.activationE == 0.0))
            c->needs_visc = true;

    hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete c;
        return set_err(SNH_EHIP, "hipStreamCreate: %s", hipGetErrorString(he));
    }
    c->own_stream = true;
    const size_t N = c->N, E = c->E;
    c->n_geo_blocks = (int)((E + GEO_BLOCK - 1) / GEO_BLOCK);
    c->n_node_blocks = (int)((N + NODE_BLOCK - 1) / NODE_BLOCK);
    bool ok = true;
#define ALLOC(p, bytes) ok = ok && (hipMalloc((void**)&(p), (bytes)) == hipSuccess)
    ALLOC(c->x[0], 3 * N * sizeof(double));
    ALLOC(c->x[1], 3 * N * sizeof(double));
    ALLOC(c->v, 3 * N * sizeof(double));
    ALLOC(c->stress, 60 * E * sizeof(double));
    if (c->needs_plastic) ALLOC(c->ps, 10 * E * sizeof(double));
    if (c->needs_visc) ALLOC(c->visc, 10 * E * sizeof(double));
    ALLOC(c->bc, N);
    ALLOC(c->part_min, (size_t)c->n_geo_blocks * sizeof(double));
    c->node_kc = 8;
    if (const char* kc = getenv("SNH_NODE_KC")) c->node_kc = atoi(kc) > 0 ? atoi(kc) : 8;
    c->n_tile_blocks = ((c->I + NT_I - 1) / NT_I) * ((c->J + NT_J - 1) / NT_J) * ((c->K + c->node_kc - 1) / c->node_kc);
    c->n_vmax_parts = c->n_node_blocks;
    ALLOC(c->part_vmax, (size_t)(c->n_node_blocks > c->n_tile_blocks ? c->n_node_blocks : c->n_tile_blocks) *
                            sizeof(double));
    ALLOC(c->scal, sizeof(DevScalars));
    ALLOC(c->dmat, SNH_MAX_MATERIALS * sizeof(DevMaterial));
    ALLOC(c->dbc, 256 * sizeof(DevBC));
    if (c->beta_dens) ALLOC(c->dens, 10 * E * sizeof(double));
    if (cfg->topo_kappa > 0) {
        const size_t P = (size_t)c->I * c->K;
        c->topo_parts = (int)((P + TOPO_PREP_BLOCK - 1) / TOPO_PREP_BLOCK);
        ALLOC(c->topo_rec, TOPO_REC * P * sizeof(double));
        ALLOC(c->topo_pmax, (size_t)c->topo_parts * sizeof(double));
        ALLOC(c->topo_pfirst, (size_t)c->topo_parts * sizeof(int));
    }
#undef ALLOC
    if (!ok) {
        free_all(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return set_err(SNH_ENOMEM, "device allocation failed (%zu nodes, %zu elements)", N, E);
    }
    if (cfg->topo_kappa > 0) {
        const int w = c->I < c->K ? c->I : c->K;
        c->topo_threads = 64;
        while (c->topo_threads < w) c->topo_threads *= 2;
    }

    /* initial mesh (BlockGeometry.c:326-349), zero velocity + BCs, zero stress */
    std::vector<double> hx(3 * N), hv(3 * N, 0.0);
    const double mn[3] = {cfg->minX, cfg->minY, cfg->minZ}, mx[3] = {cfg->maxX, cfg->maxY, cfg->maxZ};
    const int sz[3] = {c->I, c->J, c->K};
    for (size_t n = 0; n < N; n++) {
        int ijk[3] = {(int)(n % c->I), (int)((n / c->I) % c->J), (int)(n / ((size_t)c->I * c->J))};
        for (int d = 0; d < 3; d++)
            hx[d * N + n] = sz[d] > 1 ? mn[d] + (mx[d] - mn[d]) * ((double)ijk[d] / (double)(sz[d] - 1)) : mn[d];
    }
    /* BC codes: one per distinct (mask, values) combination (CompositeVC
       semantics: union of the walls, last item wins per component) */
    std::vector<uint8_t> hbc(N, 0);
    std::vector<DevBC> table(1);
    table[0].mask = 0;
    for (size_t n = 0; n < N; n++) {
        uint32_t mask = 0;
        double val[3] = {0, 0, 0};
        bool member = false;
        for (auto& b : c->bcs) {
            if (!on_wall(c, (int)n, b.wall)) continue;
            member = true;
            for (int d = 0; d < 3; d++)
                if (b.mask & (1u << d)) {
                    mask |= 1u << d;
                    val[d] = b.value[d];
                }
        }
        if (!member) continue;
        size_t q;
        for (q = 1; q < table.size(); q++)
            if (table[q].mask == mask && !memcmp(table[q].val, val, sizeof val)) break;
        if (q == table.size()) {
            if (table.size() >= 256) {
                free_all(c);
                (void)hipStreamDestroy(c->stream);
                delete c;
                return set_err(SNH_EUNSUPPORTED, "more than 255 distinct velocity conditions");
            }
            DevBC b;
            b.mask = mask;
            memcpy(b.val, val, sizeof val);
            table.push_back(b);
        }
        hbc[n] = (uint8_t)q;
        for (int d = 0; d < 3; d++)
            if (mask & (1u << d)) hv[d * N + n] = val[d];
    }
    std::vector<DevMaterial> dm(SNH_MAX_MATERIALS);
    for (size_t q = 0; q < c->mats.size(); q++) fill_dev_material(c->mats[q], dm[q]);

    int rc = SNH_OK;
    do {
        if (ctx_copy(c, c->dmat, dm.data(), sizeof(DevMaterial) * SNH_MAX_MATERIALS, hipMemcpyHostToDevice) !=
                hipSuccess ||
            ctx_copy(c, c->dbc, table.data(), sizeof(DevBC) * table.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_err(SNH_EHIP, "material / BC table upload failed");
            break;
        }
        if (ctx_copy(c, c->x[0], hx.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            ctx_copy(c, c->x[1], hx.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            ctx_copy(c, c->v, hv.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            ctx_copy(c, c->bc, hbc.data(), N, hipMemcpyHostToDevice) != hipSuccess ||
            ctx_fill(c, c->stress, 0, 60 * E * sizeof(double)) != hipSuccess ||
            (c->ps && ctx_fill(c, c->ps, 0, 10 * E * sizeof(double)) != hipSuccess) ||
            (c->visc && ctx_fill(c, c->visc, 0, 10 * E * sizeof(double)) != hipSuccess)) {
            rc = set_err(SNH_EHIP, "initial upload failed");
            break;
        }
        DevScalars s0;
        s0.dt = cfg->timeStep;
        s0.dt_rot = cfg->timeStep;
        s0.speedOfSound = 0.0;
        s0.minLen = 0.0;
        s0.vmax = 0.0;
        s0.thermal_err = 0;
        s0.nonfinite = 0;
        s0.topoGradMax = 0.0;
        if (ctx_copy(c, c->scal, &s0, sizeof s0, hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_err(SNH_EHIP, "scalar upload failed");
            break;
        }
    } while (0);
    if (rc) {
        free_all(c);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    c->cur = 0;
    c->elem_variant = SNH_ELEM_TET_PER_LANE;
    c->node_band = 8;
    c->node_variant = SNH_NODE_PER_NODE;
    if (const char* nv = getenv("SNH_NODE_VARIANT")) c->node_variant = atoi(nv) == SNH_NODE_TILED ? SNH_NODE_TILED : SNH_NODE_PER_NODE;
    if (const char* nb = getenv("SNH_NODE_BAND")) c->node_band = atoi(nb);
    c->adj = nullptr;
    c->diag_on = false;
    c->diag_buf = c->force_buf = nullptr;
    c->diag_mx = nullptr;
    c->sched_dt = c->sched_c = nullptr;
    c->sched_damp = nullptr;
    c->sched_base = 0;
    c->sched_n = 0;
    if (const char* ev = getenv("SNH_ELEM_VARIANT")) {
        const int v = atoi(ev);
        c->elem_variant = (v >= 0 && v < SNH_ELEM_VARIANTS) ? v : 0;
    }
    c->zero_done = false;
    c->rotation_pending = false;
    c->timeStep = 0;
    *out = c;
    return SNH_OK;
}

static void adj_free(struct AdjointCtx* A);

extern "C" int snh_context_destroy(snh_context* c) {
    if (!c) return SNH_OK;
    (void)hipStreamSynchronize(c->stream);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    adj_free(c->adj);
    free_all(c);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SNH_OK;
}

extern "C" int snh_set_stream(snh_context* c, void* stream) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    if (stream) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return SNH_OK;
}

extern "C" int snh_set_element_material(snh_context* c, const uint8_t* mat, size_t nelem) {
    if (!c || !mat) return set_err(SNH_EINVAL, "null argument");
    if (nelem != (size_t)c->E) return set_err(SNH_EINVAL, "expected %d elements, got %zu", c->E, nelem);
    for (size_t e = 0; e < nelem; e++)
        if (mat[e] >= c->mats.size()) return set_err(SNH_EINVAL, "element %zu: material %u out of range", e, mat[e]);
    if (!c->mat) HIPCHK(hipMalloc((void**)&c->mat, nelem));
    HIPCHK(hipMemcpyAsync(c->mat, mat, nelem, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    bool used[SNH_MAX_MATERIALS] = {};
    for (size_t e = 0; e < nelem; e++) used[mat[e]] = true;
    /* the reference's speed of sound is a running maximum over the elements
       UpdateElementMomentum has seen (Context.c:392, UpdateElement.c:62-67):
       before TimeStepZero it starts from the materials now in use, after it
       only grows */
    double cw = 0.0;
    for (size_t q = 0; q < c->mats.size(); q++)
        if (used[q]) cw = wave_speed_max(cw, c->mats[q]);
    c->c_wave = c->zero_done && c->c_wave > cw ? c->c_wave : cw;
    return SNH_OK;
}

extern "C" int snh_set_element_field(snh_context* c, int field, const double* values, size_t nelem) {
    if (!c || !values) return set_err(SNH_EINVAL, "null argument");
    if (nelem != (size_t)c->E) return set_err(SNH_EINVAL, "expected %d elements, got %zu", c->E, nelem);
    double** dst = field == SNH_FIELD_REFVISC ? &c->refvisc : field == SNH_FIELD_COHESION_OFFSET ? &c->cohoff : nullptr;
    if (!dst) return set_err(SNH_EINVAL, "unknown field %d", field);
    if (!*dst) HIPCHK(hipMalloc((void**)dst, nelem * sizeof(double)));
    HIPCHK(hipMemcpyAsync(*dst, values, nelem * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

enum { XF_VEC3 = 0, XF_TET = 1, XF_STRESS = 2 };

static int ensure_stage(snh_context* c) {
    if (!c->stage) {
        const size_t n = std::max((size_t)60 * c->E, (size_t)3 * c->N);
        HIPCHK(hipMalloc((void**)&c->stage, n * sizeof(double)));
    }
    return SNH_OK;
}

/* device rows -> host records of one field */
static int download_rows(snh_context* c, const double* rows, double* host, size_t M, int kind) {
    int rc;
    if ((rc = ensure_stage(c))) return rc;
    const dim3 grid((unsigned)((M + 63) / 64)), block(256);
    size_t C;
    switch (kind) {
    case XF_VEC3: C = 3; hipLaunchKernelGGL((k_rows_to_records<3, LAY_PLAIN>), grid, block, 0, c->stream, rows, c->stage, M); break;
    case XF_TET: C = 10; hipLaunchKernelGGL((k_rows_to_records<10, LAY_PLAIN>), grid, block, 0, c->stream, rows, c->stage, M); break;
    default: C = 60; hipLaunchKernelGGL((k_rows_to_records<60, LAY_STRESS>), grid, block, 0, c->stream, rows, c->stage, M); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(host, c->stage, C * M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

/* host records -> device rows of one field */
static int upload_rows(snh_context* c, const double* host, double* rows, size_t M, int kind) {
    int rc;
    if ((rc = ensure_stage(c))) return rc;
    const size_t C = kind == XF_VEC3 ? 3 : kind == XF_TET ? 10 : 60;
    HIPCHK(hipMemcpyAsync(c->stage, host, C * M * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const dim3 grid((unsigned)((M + 63) / 64)), block(256);
    switch (kind) {
    case XF_VEC3: hipLaunchKernelGGL((k_records_to_rows<3, LAY_PLAIN>), grid, block, 0, c->stream, c->stage, rows, M); break;
    case XF_TET: hipLaunchKernelGGL((k_records_to_rows<10, LAY_PLAIN>), grid, block, 0, c->stream, c->stage, rows, M); break;
    default: hipLaunchKernelGGL((k_records_to_rows<60, LAY_STRESS>), grid, block, 0, c->stream, c->stage, rows, M); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

static int clear_firewalls(snh_context* c);

extern "C" int snh_upload_state(snh_context* c, const snh_state* s) {
    Range range_("snh_upload_state");
    if (!c || !s) return set_err(SNH_EINVAL, "null argument");
    const size_t N = c->N, E = c->E;
    int rc;
    if (s->temperature && !c->thermal)
        return set_err(SNH_EINVAL, "no temperature plugin in this context (snh_set_thermal)");
    if (s->plastic_strain && !c->ps) return set_err(SNH_EINVAL, "no plastic rheology in this context");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (s->coord) {
        if ((rc = upload_rows(c, s->coord, c->x[c->cur], N, XF_VEC3))) return rc;
        /* on the context's stream: a device-to-device hipMemcpy may still run
           when the host returns, unordered with that (non-blocking) stream */
        HIPCHK(hipMemcpyAsync(c->x[c->cur ^ 1], c->x[c->cur], 3 * N * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (s->velocity && (rc = upload_rows(c, s->velocity, c->v, N, XF_VEC3))) return rc;
    if (s->stress) {
        if ((rc = upload_rows(c, s->stress, c->stress, E, XF_STRESS))) return rc;
        c->rotation_pending = false;
    }
    if ((s->coord || s->velocity || s->stress) && c->zero_done) {
        /* uploaded arrays are the state after LoopElementsMomentum (stress
           already rotated): cancel the pending rotation */
        const double zero = 0.0;
        HIPCHK(ctx_copy(c, (char*)c->scal + offsetof(DevScalars, dt_rot), &zero, sizeof zero, hipMemcpyHostToDevice));
        c->rotation_pending = false;
    }
    if (s->temperature)
        HIPCHK(ctx_copy(c, c->temp, s->temperature, (size_t)c->N * sizeof(double), hipMemcpyHostToDevice));
    if (s->plastic_strain && (rc = upload_rows(c, s->plastic_strain, c->ps, E, XF_TET))) return rc;
    return clear_firewalls(c);
}

static int launch_geo(snh_context* c) {
    KTimer kt(c, KK_GEO);
    StepArgs a = make_args(c);
    hipLaunchKernelGGL(k_geo, dim3(c->n_geo_blocks), dim3(GEO_BLOCK), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

static int launch_fin(snh_context* c) {
    KTimer kt(c, KK_FIN);
    StepArgs a = make_args(c);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(FIN_BLOCK), 0, c->stream, a, c->n_geo_blocks, c->n_vmax_parts);
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

static int launch_elem_args(snh_context* c, StepArgs a);
static int launch_elem_fp64(snh_context* c, StepArgs a);
static int launch_elem(snh_context* c) {
    int rc;
    if (c->cfg.precision == SNH_PRECISION_MIXED && c->diag_on) {
        /* the dump's element quantities of a mixed-precision step: the FP64
           element pass on the same state, its stress and plastic strain
           written to scratch, before the mixed pass updates the state */
        StepArgs d = make_args(c);
        d.stress_out = c->diag_mx;
        d.ps_out = c->ps ? c->diag_mx + 60 * (size_t)c->E : nullptr;
        d.visc = nullptr;
        d.force_out = nullptr;
        if ((rc = launch_elem_fp64(c, d))) return rc;
    }
    {
        KTimer kt(c, KK_ELEM);
        rc = launch_elem_args(c, make_args(c));
    }
    if (rc == SNH_OK && c->beta_dens) {
        /* the pressure-dependent tet densities of this Stress call */
        StepArgs a = make_args(c);
        const dim3 g(c->n_geo_blocks), b(GEO_BLOCK);
        if (a.mat && a.tpre) hipLaunchKernelGGL((k_density<true, true>), g, b, 0, c->stream, a);
        else if (a.mat) hipLaunchKernelGGL((k_density<true, false>), g, b, 0, c->stream, a);
        else if (a.tpre) hipLaunchKernelGGL((k_density<false, true>), g, b, 0, c->stream, a);
        else hipLaunchKernelGGL((k_density<false, false>), g, b, 0, c->stream, a);
        HIPCHK(hipGetLastError());
    }
    return rc;
}
static int launch_elem_fp64(snh_context* c, StepArgs a);

static int launch_elem_args(snh_context* c, StepArgs a) {
    dim3 grid((c->E + EPW - 1) / EPW), block(64);
    if (c->cfg.precision == SNH_PRECISION_MIXED) {
        /* tet per wave unless the tet-per-lane layout is forced */
        const bool tw = c->elem_variant != SNH_ELEM_TET_PER_LANE;
        if (tw) {
            grid = dim3((c->E + 63) / 64);
            block = dim3(640);
        }
#define MX_LAUNCH(R)                                                                          \
    do {                                                                                      \
        if (tw) hipLaunchKernelGGL((k_elem_mx<R, true>), grid, block, 0, c->stream, a);       \
        else hipLaunchKernelGGL((k_elem_mx<R, false>), grid, block, 0, c->stream, a);         \
    } while (0)
        switch (c->rset) {
        case RS_ELASTIC: MX_LAUNCH(RS_ELASTIC); break;
        case RS_PLASTIC: MX_LAUNCH(RS_PLASTIC); break;
        case RS_MAXWELL: MX_LAUNCH(RS_MAXWELL); break;
        default: MX_LAUNCH(RS_ELASTIC | RS_PLASTIC | RS_MAXWELL); break;
        }
#undef MX_LAUNCH
        HIPCHK(hipGetLastError());
        return SNH_OK;
    }
    return launch_elem_fp64(c, a);
}

/* the FP64 element pass in the context's layout */
static int launch_elem_fp64(snh_context* c, StepArgs a) {
    dim3 grid((c->E + EPW - 1) / EPW), block(64);
    const bool mm = a.mat != nullptr;
    /* the diagnostics are written by the element-per-lane variant only */
    if (a.diag || c->elem_variant == SNH_ELEM_ELEMENT_PER_LANE ||
        c->elem_variant == SNH_ELEM_ELEMENT_PER_LANE_LOWREG) {
        dim3 g2((c->E + ELEM_BLOCK - 1) / ELEM_BLOCK), b2(ELEM_BLOCK);
        const bool lr = !a.diag && c->elem_variant == SNH_ELEM_ELEMENT_PER_LANE_LOWREG;
#define SEQ_LAUNCH(R)                                                                         \
    do {                                                                                      \
        if (mm) {                                                                             \
            if (lr) hipLaunchKernelGGL((k_elem_seq<R, true, true>), g2, b2, 0, c->stream, a); \
            else hipLaunchKernelGGL((k_elem_seq<R, false, true>), g2, b2, 0, c->stream, a);   \
        } else {                                                                              \
            if (lr) hipLaunchKernelGGL((k_elem_seq<R, true, false>), g2, b2, 0, c->stream, a); \
            else hipLaunchKernelGGL((k_elem_seq<R, false, false>), g2, b2, 0, c->stream, a);  \
        }                                                                                     \
    } while (0)
        switch (c->rset) {
        case RS_ELASTIC: SEQ_LAUNCH(RS_ELASTIC); break;
        case RS_PLASTIC: SEQ_LAUNCH(RS_PLASTIC); break;
        case RS_MAXWELL: SEQ_LAUNCH(RS_MAXWELL); break;
        case RS_VISCOPLASTIC: SEQ_LAUNCH(RS_VISCOPLASTIC); break;
        default: SEQ_LAUNCH(RS_ELASTIC | RS_PLASTIC | RS_MAXWELL | RS_VISCOPLASTIC); break;
        }
#undef SEQ_LAUNCH
        HIPCHK(hipGetLastError());
        return SNH_OK;
    }
    const bool tw = c->elem_variant == SNH_ELEM_TET_PER_WAVE;
    if (tw) {
        grid = dim3((c->E + 63) / 64);
        block = dim3(640);
    }
#define TL_LAUNCH(R)                                                                          \
    do {                                                                                      \
        if (mm) {                                                                             \
            if (tw) hipLaunchKernelGGL((k_elem<R, true, true>), grid, block, 0, c->stream, a); \
            else hipLaunchKernelGGL((k_elem<R, false, true>), grid, block, 0, c->stream, a);  \
        } else {                                                                              \
            if (tw) hipLaunchKernelGGL((k_elem<R, true, false>), grid, block, 0, c->stream, a); \
            else hipLaunchKernelGGL((k_elem<R, false, false>), grid, block, 0, c->stream, a); \
        }                                                                                     \
    } while (0)
    switch (c->rset) {
    case RS_ELASTIC: TL_LAUNCH(RS_ELASTIC); break;
    case RS_PLASTIC: TL_LAUNCH(RS_PLASTIC); break;
    case RS_MAXWELL: TL_LAUNCH(RS_MAXWELL); break;
    case RS_VISCOPLASTIC: TL_LAUNCH(RS_VISCOPLASTIC); break;
    default: TL_LAUNCH(RS_ELASTIC | RS_PLASTIC | RS_MAXWELL | RS_VISCOPLASTIC); break;
    }
#undef TL_LAUNCH
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

/* index of timeStep t in the prescribed schedule, or -1 */
static int sched_index(const snh_context* c, uint32_t t) {
    if (c->sched_n <= 0 || t < c->sched_base || t - c->sched_base >= (uint32_t)c->sched_n) return -1;
    return (int)(t - c->sched_base);
}

/* the FP64 node pass in the context's variant; records how many block
   maxima of |v| it leaves for k_fin */
static void launch_node_fp64(snh_context* c, const StepArgs& a) {
    if (c->node_variant == SNH_NODE_TILED) {
        const dim3 g(c->n_tile_blocks), b(NT_THREADS);
        if (a.dens && a.mat) hipLaunchKernelGGL((k_node_tile<true, true>), g, b, 0, c->stream, a);
        else if (a.dens) hipLaunchKernelGGL((k_node_tile<false, true>), g, b, 0, c->stream, a);
        else if (a.mat) hipLaunchKernelGGL((k_node_tile<true, false>), g, b, 0, c->stream, a);
        else hipLaunchKernelGGL((k_node_tile<false, false>), g, b, 0, c->stream, a);
        c->n_vmax_parts = c->n_tile_blocks;
        return;
    }
    const dim3 g(c->n_node_blocks), b(NODE_BLOCK);
    if (a.dens && a.mat) hipLaunchKernelGGL((k_node<true, true>), g, b, 0, c->stream, a);
    else if (a.dens) hipLaunchKernelGGL((k_node<false, true>), g, b, 0, c->stream, a);
    else if (a.mat) hipLaunchKernelGGL((k_node<true, false>), g, b, 0, c->stream, a);
    else hipLaunchKernelGGL((k_node<false, false>), g, b, 0, c->stream, a);
    c->n_vmax_parts = c->n_node_blocks;
}

/* refStep: the reference's timeStep of the step being run (1 for the first) */
static int launch_node(snh_context* c, int refStep) {
    KTimer kt(c, KK_NODE);
    StepArgs a = make_args(c);
    const int si = sched_index(c, c->timeStep);
    if (si >= 0 && c->sched_damp) a.damp_sched = c->sched_damp + (size_t)si * c->N;
    if (c->recording) a.tape_code = c->rec_code;
    if (c->cfg.precision == SNH_PRECISION_MIXED) {
        hipLaunchKernelGGL(k_node_mx, dim3(c->n_node_blocks), dim3(NODE_BLOCK), 0, c->stream, a);
        c->n_vmax_parts = c->n_node_blocks;
    } else {
        launch_node_fp64(c, a);
    }
    HIPCHK(hipGetLastError());
    if (c->cfg.topo_kappa > 0 && refStep > 1 && topo_mode(c->I, c->K) != TOPO_NONE) {
        /* before the buffer flip: xc still holds the old coordinates */
        TopoArgs t;
        t.mode = topo_mode(c->I, c->K);
        t.rec = c->topo_rec;
        t.pmax = c->topo_pmax;
        t.pfirst = c->topo_pfirst;
        t.nparts = c->topo_parts;
        t.kappa = c->cfg.topo_kappa;
        t.crit = c->cfg.topoGradCriterion;
        hipLaunchKernelGGL(k_topo_prep, dim3(c->topo_parts), dim3(TOPO_PREP_BLOCK), 0, c->stream, a, t);
        HIPCHK(hipGetLastError());
        if (t.mode == TOPO_ALONG_X)
            hipLaunchKernelGGL((k_topo_line<TOPO_ALONG_X, 16>), dim3(1), dim3(64), 0, c->stream, a, t);
        else if (t.mode == TOPO_ALONG_Z)
            hipLaunchKernelGGL((k_topo_line<TOPO_ALONG_Z, 16>), dim3(1), dim3(64), 0, c->stream, a, t);
        else if (c->topo_threads <= 256)
            hipLaunchKernelGGL((k_topo<256, 8>), dim3(1), dim3(c->topo_threads), 0, c->stream, a, t);
        else
            hipLaunchKernelGGL((k_topo<1024, 3>), dim3(1), dim3(c->topo_threads), 0, c->stream, a, t);
        HIPCHK(hipGetLastError());
    }
    c->cur ^= 1;
    return SNH_OK;
}

static int launch_heat(snh_context* c, bool elem) {
    StepArgs a = make_args(c);
    if (elem) {
        KTimer kt(c, KK_HEAT_ELEM);
        if (a.mat)
            hipLaunchKernelGGL(k_heat_elem<true>, dim3(c->n_geo_blocks), dim3(GEO_BLOCK), 0, c->stream, a);
        else
            hipLaunchKernelGGL(k_heat_elem<false>, dim3(c->n_geo_blocks), dim3(GEO_BLOCK), 0, c->stream, a);
    } else {
        KTimer kt(c, KK_HEAT_NODE);
        if (a.mat)
            hipLaunchKernelGGL(k_heat_node<true>, dim3(c->n_node_blocks), dim3(NODE_BLOCK), 0, c->stream, a);
        else
            hipLaunchKernelGGL(k_heat_node<false>, dim3(c->n_node_blocks), dim3(NODE_BLOCK), 0, c->stream, a);
    }
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

/* tet density phsDensity of every tet, as TimeStepZero sets it
   (Context.c:1009-1017) */
static int fill_tet_density(snh_context* c) {
    std::vector<uint8_t> mat(c->E, 0);
    if (c->mat) HIPCHK(ctx_copy(c, mat.data(), c->mat, c->E, hipMemcpyDeviceToHost));
    std::vector<double> d(10 * (size_t)c->E);
    for (size_t e = 0; e < (size_t)c->E; e++)
        for (int t = 0; t < NTET; t++) d[(size_t)t * c->E + e] = c->cfg.materials[mat[e]].phsDensity;
    HIPCHK(ctx_copy(c, c->dens, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    return SNH_OK;
}

extern "C" int snh_set_thermal(snh_context* c, const snh_thermal* th) {
    if (!c || !th) return set_err(SNH_EINVAL, "null argument");
    if (th->nbcs && !th->bcs) return set_err(SNH_EINVAL, "null temperature BC list");
    if (c->zero_done) return set_err(SNH_ESTATE, "snh_set_thermal after snh_timestep_zero");
    if (c->adj) return set_err(SNH_ESTATE, "snh_set_thermal during an adjoint recording");
    for (uint32_t k = 0; k < c->cfg.nmaterials; k++) {
        const snh_material& m = c->cfg.materials[k];
        if (!(m.heatCapacity > 0.0) || !(m.phsDensity > 0.0) || !(m.thermal_conduct >= 0.0))
            return set_err(SNH_EINVAL, "temperature plugin needs heatCapacity > 0, phsDensity > 0, thermal_conduct >= 0");
    }
    const size_t N = c->N, E = c->E;
    if (!c->temp) {
        HIPCHK(hipMalloc((void**)&c->temp, N * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c->temp0, N * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c->tpre, N * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c->hflux, 30 * E * sizeof(double)));
        if (!c->dens) HIPCHK(hipMalloc((void**)&c->dens, 10 * E * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c->tbc, N));
        HIPCHK(hipMalloc((void**)&c->tbcv, 256 * sizeof(double)));
        HIPCHK(ctx_fill(c, c->hflux, 0, 30 * E * sizeof(double)));
    }
    if (th->computeThermalStress && !c->tstress) {
        HIPCHK(hipMalloc((void**)&c->tstress, 10 * E * sizeof(double)));
        HIPCHK(ctx_fill(c, c->tstress, 0, 10 * E * sizeof(double)));
    }
    /* _SnacTemperature_InitialConditions (VariableConditions.c:155-212):
       T0 = T, then the temperatureBCs (later entries win) */
    std::vector<double> T(N, 0.0), T0;
    if (th->temperature) T.assign(th->temperature, th->temperature + N);
    T0 = T;
    std::vector<uint8_t> code(N, 0);
    std::vector<double> vals(256, 0.0);
    int nvals = 1;
    for (size_t n = 0; n < N; n++) {
        int b_last = -1;
        for (uint32_t b = 0; b < th->nbcs; b++)
            if (on_wall(c, (int)n, th->bcs[b].wall)) b_last = (int)b;
        if (b_last < 0) continue;
        const double v = th->bcs[b_last].value;
        int k = 1;
        while (k < nvals && memcmp(&vals[k], &v, sizeof v) != 0) k++;
        if (k == nvals) {
            if (nvals == 256) return set_err(SNH_EINVAL, "more than 255 distinct temperature BC values");
            vals[nvals++] = v;
        }
        code[n] = (uint8_t)k;
        T[n] = v;
    }
    HIPCHK(ctx_copy(c, c->temp, T.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx_copy(c, c->temp0, T0.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx_copy(c, c->tpre, T.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx_copy(c, c->tbc, code.data(), N, hipMemcpyHostToDevice));
    HIPCHK(ctx_copy(c, c->tbcv, vals.data(), 256 * sizeof(double), hipMemcpyHostToDevice));
    c->thermal = true;
    c->thermal_stress = th->computeThermalStress != 0;
    return fill_tet_density(c);
}

extern "C" int snh_remesh(snh_context* c, const snh_state* s) {
    Range range_("snh_remesh");
    if (!c || !s) return set_err(SNH_EINVAL, "null argument");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_remesh before snh_timestep_zero");
    if (c->adj) return set_err(SNH_ESTATE, "snh_remesh during an adjoint recording");
    /* a mixed-precision context holds the same FP64 state; its pending
       rotation is applied here by the FP64 kernel */
    int rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t N = c->N, E = c->E;
    /* the Jaumann rotation of the last LoopElementsMomentum, on the old mesh */
    if (c->rotation_pending) {
        if (!c->tmp) HIPCHK(hipMalloc((void**)&c->tmp, 60 * E * sizeof(double)));
        StepArgs a = make_args(c);
        hipLaunchKernelGGL(k_rotated, dim3((c->E + ELEM_BLOCK - 1) / ELEM_BLOCK), dim3(ELEM_BLOCK), 0, c->stream, a,
                           c->tmp);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->stress, c->tmp, 60 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->rotation_pending = false;
    }
    /* the remeshed arrays; the coordinates before the remesh become the
       previous ones, whose tet volumes are the old_volume of the next step */
    if (s->coord) {
        HIPCHK(hipMemcpyAsync(c->x[c->cur ^ 1], c->x[c->cur], 3 * N * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
        if ((rc = upload_rows(c, s->coord, c->x[c->cur], N, XF_VEC3))) return rc;
    }
    snh_state rest = *s;
    rest.coord = nullptr;
    rest.viscosity = nullptr;
    rest.volume = nullptr;
    if ((rc = snh_upload_state(c, &rest))) return rc;
    /* _SnacRemesher_UpdateElements: geometry of the new mesh, its rotation
       with the dt of the last Sync (applied by the next CalcStresses), and the
       running minimum of the length scale */
    if ((rc = launch_geo(c))) return rc;
    StepArgs a = make_args(c);
    hipLaunchKernelGGL(k_minlen_merge, dim3(1), dim3(FIN_BLOCK), 0, c->stream, a, c->n_geo_blocks);
    HIPCHK(hipGetLastError());
    c->rotation_pending = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

extern "C" int snh_loop_elements_energy(snh_context* c) {
    Range range_("snh_loop_elements_energy");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->thermal) return set_err(SNH_ESTATE, "no temperature plugin (snh_set_thermal)");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_loop_elements_energy before snh_timestep_zero");
    return launch_heat(c, true);
}

extern "C" int snh_loop_nodes_energy(snh_context* c) {
    Range range_("snh_loop_nodes_energy");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->thermal) return set_err(SNH_ESTATE, "no temperature plugin (snh_set_thermal)");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_loop_nodes_energy before snh_timestep_zero");
    return launch_heat(c, false);
}

extern "C" int snh_hydrostatic_ic(snh_context* c) {
    Range range_("snh_hydrostatic_ic");
    if (!c) return set_err(SNH_EINVAL, "null context");
    const size_t N = c->N, E = c->E;
    std::vector<double> x(3 * N), st(60 * E);
    std::vector<uint8_t> mat;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(ctx_copy(c, x.data(), c->x[c->cur], 3 * N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(ctx_copy(c, st.data(), c->stress, 60 * E * sizeof(double), hipMemcpyDeviceToHost));
    if (c->mat) {
        mat.resize(E);
        HIPCHK(ctx_copy(c, mat.data(), c->mat, E, hipMemcpyDeviceToHost));
    }
    const int elx = c->EI, ely = c->EJ, elz = c->EK;
    auto yof = [&](int i, int j, int k) { return x[N + (size_t)(i + c->I * (j + c->J * k))]; };
    for (int elK = 0; elK < elz; elK++)
        for (int elI = 0; elI < elx; elI++) {
            double rogh = 0.0;
            for (int elJ = ely - 1; elJ > -1; elJ--) {
                const size_t e = (size_t)elI + (size_t)elJ * elx + (size_t)elK * elx * ely;
                const snh_material& m = c->mats[c->mat ? mat[e] : 0];
                double densT = 0.0, dhE[4], dh, dPT, dP, P;
                /* tet densities are still zero before TimeStepZero */
                for (int t = 0; t < NTET; t++) densT += (m.phsDensity * (1.0 - m.alpha * (0.0 - m.reftemp))) / NTET;
                /* element nodes 6-0, 2-4, 3-5, 7-1 (HexaEL.c:462-469 numbering) */
                dhE[0] = yof(elI + 1, elJ + 1, elK + 1) - yof(elI, elJ, elK);
                dhE[1] = yof(elI + 1, elJ + 1, elK) - yof(elI, elJ, elK + 1);
                dhE[2] = yof(elI, elJ + 1, elK) - yof(elI + 1, elJ, elK + 1);
                dhE[3] = yof(elI, elJ + 1, elK + 1) - yof(elI + 1, elJ, elK);
                dh = 0.25 * (dhE[0] + dhE[1] + dhE[2] + dhE[3]);
                dPT = densT * c->cfg.gravity * dh;
                dP = dPT * (1.0 - m.beta * rogh) / (1.0 + m.beta / 2.0 * dPT);
                P = rogh + 0.5 * dP;
                P = -1.0 * (P + 0.0);
                for (int t = 0; t < NTET; t++)
                    for (int q = 0; q < 3; q++) st[((size_t)q * NTET + t) * E + e] = P;
                rogh += dP;
            }
        }
    HIPCHK(ctx_copy(c, c->stress, st.data(), 60 * E * sizeof(double), hipMemcpyHostToDevice));
    c->rotation_pending = false;
    return SNH_OK;
}

extern "C" int snh_timestep_zero(snh_context* c) {
    Range range_("snh_timestep_zero");
    if (!c) return set_err(SNH_EINVAL, "null context");
    int rc;
    /* a run (or a restart, a new run of the reference) begins with the
       running topography gradient at zero (Context.c:759-764) */
    const double zero = 0.0;
    HIPCHK(hipMemcpyAsync((char*)c->scal + offsetof(DevScalars, topoGradMax), &zero, sizeof zero,
                          hipMemcpyHostToDevice, c->stream));
    StepArgs a = make_args(c);
    hipLaunchKernelGGL(k_vmax, dim3(c->n_node_blocks), dim3(NODE_BLOCK), 0, c->stream, a);
    c->n_vmax_parts = c->n_node_blocks;
    HIPCHK(hipGetLastError());
    if ((rc = launch_geo(c))) return rc;
    if ((rc = launch_fin(c))) return rc;
    if (c->thermal && (rc = fill_tet_density(c))) return rc;
    c->zero_done = true;
    c->rotation_pending = true;
    c->timeStep = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

extern "C" int snh_calc_stresses(snh_context* c) {
    Range range_("snh_calc_stresses");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_calc_stresses before snh_timestep_zero");
    int rc = launch_elem(c);
    if (rc) return rc;
    c->rotation_pending = false;
    c->timeStep++;
    return SNH_OK;
}

extern "C" int snh_loop_nodes_momentum(snh_context* c) {
    Range range_("snh_loop_nodes_momentum");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_loop_nodes_momentum before snh_timestep_zero");
    return launch_node(c, (int)c->timeStep);   /* snh_calc_stresses counted this step */
}

extern "C" int snh_loop_elements_momentum(snh_context* c) {
    Range range_("snh_loop_elements_momentum");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_loop_elements_momentum before snh_timestep_zero");
    int rc = launch_geo(c);
    if (rc) return rc;
    c->rotation_pending = true;
    return SNH_OK;
}

extern "C" int snh_sync(snh_context* c) {
    Range range_("snh_sync");
    if (!c) return set_err(SNH_EINVAL, "null context");
    return launch_fin(c);
}

extern "C" int snh_step(snh_context* c, int nsteps) {
    Range range_("snh_step");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_step before snh_timestep_zero");
    if (nsteps < 0) return set_err(SNH_EINVAL, "negative step count");
    for (int s = 0; s < nsteps; s++) {
        int rc;
        const int si = sched_index(c, c->timeStep);
        if (si >= 0) {
            hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, c->stream, c->scal, c->sched_dt + si,
                               c->sched_c + si);
            HIPCHK(hipGetLastError());
        }
        if (c->thermal) {   /* _Snac_Context_Solve, Context.c:1073-1074 */
            if ((rc = launch_heat(c, true))) return rc;
            if ((rc = launch_heat(c, false))) return rc;
        }
        if ((rc = launch_elem(c))) return rc;
        if ((rc = launch_node(c, (int)c->timeStep + 1))) return rc;
        if ((rc = launch_geo(c))) return rc;
        if ((rc = launch_fin(c))) return rc;
        c->timeStep++;
    }
    c->rotation_pending = c->rotation_pending || nsteps > 0;
    return SNH_OK;
}

extern "C" int snh_set_diagnostics(snh_context* c, int enable) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (enable && c->cfg.precision == SNH_PRECISION_MIXED && !c->diag_mx)
        HIPCHK(hipMalloc((void**)&c->diag_mx, 70 * (size_t)c->E * sizeof(double)));
    if (enable && !c->diag_buf) {
        HIPCHK(hipMalloc((void**)&c->diag_buf, 5 * (size_t)c->E * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c->force_buf, 3 * (size_t)c->N * sizeof(double)));
        HIPCHK(ctx_fill(c, c->diag_buf, 0, 5 * (size_t)c->E * sizeof(double)));
        HIPCHK(ctx_fill(c, c->force_buf, 0, 3 * (size_t)c->N * sizeof(double)));
    }
    c->diag_on = enable != 0;
    return SNH_OK;
}

extern "C" int snh_get_diagnostics(snh_context* c, const snh_diagnostics* out) {
    if (!c || !out) return set_err(SNH_EINVAL, "null argument");
    if (!c->diag_buf) return set_err(SNH_ESTATE, "no diagnostics recorded (snh_set_diagnostics)");
    const size_t E = c->E, N = c->N;
    HIPCHK(hipStreamSynchronize(c->stream));
    double* dst[4] = {out->strain_rate, out->stress, out->hydro_pressure, out->plastic_strain};
    for (int k = 0; k < 4; k++)
        if (dst[k]) HIPCHK(ctx_copy(c, dst[k], c->diag_buf + k * E, E * sizeof(double), hipMemcpyDeviceToHost));
    if (out->force) {
        std::vector<double> f(3 * N);
        HIPCHK(ctx_copy(c, f.data(), c->force_buf, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t n = 0; n < N; n++)
            for (int d = 0; d < 3; d++) out->force[3 * n + d] = f[d * N + n];
    }
    return SNH_OK;
}

extern "C" int snh_set_schedule(snh_context* c, int32_t nsteps, const double* dt, const double* cs,
                                const int8_t* damp) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (nsteps < 0 || (nsteps > 0 && (!dt || !cs))) return set_err(SNH_EINVAL, "bad schedule");
    HIPCHK(hipStreamSynchronize(c->stream));
    void* old[] = {c->sched_dt, c->sched_c, c->sched_damp};
    for (void* p : old)
        if (p) (void)hipFree(p);
    c->sched_dt = c->sched_c = nullptr;
    c->sched_damp = nullptr;
    c->sched_n = 0;
    if (nsteps == 0) return SNH_OK;
    const size_t N = c->N;
    HIPCHK(hipMalloc((void**)&c->sched_dt, nsteps * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c->sched_c, nsteps * sizeof(double)));
    HIPCHK(ctx_copy(c, c->sched_dt, dt, nsteps * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(ctx_copy(c, c->sched_c, cs, nsteps * sizeof(double), hipMemcpyHostToDevice));
    if (damp) {
        std::vector<uint8_t> packed((size_t)nsteps * N);
        for (size_t i = 0; i < packed.size(); i++) {
            int code = 0;
            for (int d = 0; d < 3; d++) {
                const int v = damp[3 * i + d];
                if (v < -1 || v > 1) return set_err(SNH_EINVAL, "damping codes must be -1, 0 or 1");
                code |= (v + 1) << (2 * d);
            }
            packed[i] = (uint8_t)code;
        }
        HIPCHK(hipMalloc((void**)&c->sched_damp, packed.size()));
        HIPCHK(ctx_copy(c, c->sched_damp, packed.data(), packed.size(), hipMemcpyHostToDevice));
    }
    c->sched_base = c->timeStep;
    c->sched_n = nsteps;
    return SNH_OK;
}

/* device firewall flags (non-finite velocity or dt; the thermal dt bound)
   back to zero: done once a raised flag has been reported, and by every
   state upload, so that a caller who restarts from a checkpoint runs on.
   The context's stream is idle at both call sites. */
extern "C" int snh_record_schedule(snh_context* c, int32_t nsteps, double* dt, double* cs, int8_t* damp) {
    Range range_("snh_record_schedule");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (nsteps < 0 || (nsteps > 0 && (!dt || !cs))) return set_err(SNH_EINVAL, "bad arguments");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_record_schedule before snh_timestep_zero");
    if (c->adj) return set_err(SNH_ESTATE, "snh_record_schedule while an adjoint record is open");
    const size_t N = c->N;
    std::vector<uint8_t> codes(damp ? N : 0);
    if (damp && !c->rec_code) HIPCHK(hipMalloc((void**)&c->rec_code, N));
    uint8_t* rec = c->rec_code;
    int rc = SNH_OK;
    for (int32_t s = 0; s < nsteps && rc == SNH_OK; s++) {
        DevScalars sc;
        /* dt and speed of sound the step runs with (a schedule, if one is
           set, puts its own in place first, as snh_step does) */
        const int si = sched_index(c, c->timeStep);
        if (si >= 0) {
            hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, c->stream, c->scal, c->sched_dt + si,
                               c->sched_c + si);
            if (hipGetLastError() != hipSuccess) {
                rc = set_err(SNH_EHIP, "schedule launch failed");
                break;
            }
        }
        if (hipMemcpyAsync(&sc, c->scal, sizeof sc, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = set_err(SNH_EHIP, "scalar download failed");
            break;
        }
        dt[s] = sc.dt;
        cs[s] = sc.speedOfSound;
        c->recording = damp != nullptr;
        rc = snh_step(c, 1);
        c->recording = false;
        if (rc) break;
        if (damp) {
            if (hipMemcpyAsync(codes.data(), rec, N, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = set_err(SNH_EHIP, "damping record download failed");
                break;
            }
            int8_t* out = damp + (size_t)s * 3 * N;
            for (size_t n = 0; n < N; n++)
                for (int d = 0; d < 3; d++) out[3 * n + d] = (int8_t)((int)((codes[n] >> (2 * d)) & 3u) - 1);
        }
    }
    return rc;   /* rec_code is kept for the next recording and freed with the context */
}

static int clear_firewalls(snh_context* c) {
    const int32_t zero[2] = {0, 0};
    HIPCHK(hipMemcpyAsync((char*)c->scal + offsetof(DevScalars, thermal_err), zero, sizeof zero,
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

/* the release-mode stress firewall (Stress.c:161-166) as a non-finite
   velocity or dt, and the Journal_Firewall of Snac_Heat
   (temperature/Heat.c:121-126), raised by any step since the last check;
   each raised flag is reported once */
static int check_firewalls(snh_context* c) {
    int32_t flags[2] = {0, 0};
    HIPCHK(ctx_copy(c, flags, (char*)c->scal + offsetof(DevScalars, thermal_err), sizeof flags, hipMemcpyDeviceToHost));
    if (!flags[0] && !flags[1]) return SNH_OK;
    int rc;
    if ((rc = clear_firewalls(c))) return rc;
    if (flags[1]) return set_err(SNH_ENUMERIC, "timeStep %u: a velocity or dt is not finite (a NaN or Inf "
                                               "reached the state; Stress.c:161-166)", c->timeStep);
    if (c->thermal)
        return set_err(SNH_ENUMERIC, "dt_thermal / dt_mech <= 1: check lambda, mu, density, minLengthScale, "
                                     "thermal_conduct, heatCapacity (temperature/Heat.c:121-126)");
    return SNH_OK;
}

extern "C" int snh_synchronize(snh_context* c) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return check_firewalls(c);
}

extern "C" int snh_get_scalars(snh_context* c, snh_scalars* out) {
    if (!c || !out) return set_err(SNH_EINVAL, "null argument");
    DevScalars s;
    HIPCHK(hipMemcpyAsync(&s, c->scal, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    out->dt = s.dt;
    out->speedOfSound = s.speedOfSound;
    out->minLengthScale = s.minLen;
    out->vmax = s.vmax;
    out->timeStep = c->timeStep;
    out->topoGradMax = s.topoGradMax;
    return SNH_OK;
}

extern "C" int snh_download_state(snh_context* c, snh_state* s) {
    Range range_("snh_download_state");
    if (!c || !s) return set_err(SNH_EINVAL, "null argument");
    const size_t N = c->N, E = c->E;
    int rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (s->temperature && !c->thermal)
        return set_err(SNH_EINVAL, "no temperature plugin in this context (snh_set_thermal)");
    if (s->coord && (rc = download_rows(c, c->x[c->cur], s->coord, N, XF_VEC3))) return rc;
    if (s->temperature) HIPCHK(ctx_copy(c, s->temperature, c->temp, N * sizeof(double), hipMemcpyDeviceToHost));
    if (s->velocity && (rc = download_rows(c, c->v, s->velocity, N, XF_VEC3))) return rc;
    if (s->stress || s->volume) {
        if (!c->tmp) HIPCHK(hipMalloc((void**)&c->tmp, 60 * E * sizeof(double)));
    }
    if (s->stress) {
        const double* src = c->stress;
        if (c->rotation_pending) {
            StepArgs a = make_args(c);
            hipLaunchKernelGGL(k_rotated, dim3((c->E + ELEM_BLOCK - 1) / ELEM_BLOCK), dim3(ELEM_BLOCK), 0, c->stream,
                               a, c->tmp);
            HIPCHK(hipGetLastError());
            src = c->tmp;
        }
        if ((rc = download_rows(c, src, s->stress, E, XF_STRESS))) return rc;
    }
    if (s->volume) {
        StepArgs a = make_args(c);
        hipLaunchKernelGGL(k_volumes, dim3((c->E + 127) / 128), dim3(128), 0, c->stream, a, c->tmp);
        HIPCHK(hipGetLastError());
        if ((rc = download_rows(c, c->tmp, s->volume, E, XF_TET))) return rc;
    }
    if (s->plastic_strain) {
        if (c->ps) {
            if ((rc = download_rows(c, c->ps, s->plastic_strain, E, XF_TET))) return rc;
        } else {
            memset(s->plastic_strain, 0, 10 * E * sizeof(double));
        }
    }
    if (s->viscosity) {
        if (c->visc) {
            if ((rc = download_rows(c, c->visc, s->viscosity, E, XF_TET))) return rc;
        } else {
            /* linear law: the viscosity is the clamped reference viscosity */
            std::vector<double> rv;
            if (c->refvisc) {
                rv.resize(E);
                HIPCHK(ctx_copy(c, rv.data(), c->refvisc, E * sizeof(double), hipMemcpyDeviceToHost));
            }
            std::vector<uint8_t> mat;
            if (c->mat) {
                mat.resize(E);
                HIPCHK(ctx_copy(c, mat.data(), c->mat, E, hipMemcpyDeviceToHost));
            }
            for (size_t e = 0; e < E; e++) {
                const snh_material& m = c->mats[c->mat ? mat[e] : 0];
                double vis = 0.0;
                if ((m.rheology & (SNH_RHEOLOGY_MAXWELL | SNH_RHEOLOGY_VISCOPLASTIC)) && c->timeStep > 0) {
                    vis = c->refvisc ? rv[e] : m.refvisc;
                    if (vis < m.vis_min) vis = m.vis_min;
                    if (vis > m.vis_max) vis = m.vis_max;
                }
                for (int t = 0; t < NTET; t++) s->viscosity[e * NTET + t] = vis;
            }
        }
    }
    return SNH_OK;
}

extern "C" int snh_profile(snh_context* c, int enable) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    c->recs.clear();
    c->ev_used = 0;
    c->profiling = enable != 0;
    return SNH_OK;
}

extern "C" int snh_get_kernel_times(snh_context* c, snh_kernel_times* out) {
    if (!c || !out) return set_err(SNH_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    double ms[KK_KINDS] = {0, 0, 0, 0, 0, 0};
    uint32_t n[KK_KINDS] = {0, 0, 0, 0, 0, 0};
    for (auto& r : c->recs) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
        ms[r.kind] += t;
        n[r.kind]++;
    }
    out->elem_ms = ms[KK_ELEM]; out->node_ms = ms[KK_NODE]; out->geo_ms = ms[KK_GEO]; out->fin_ms = ms[KK_FIN];
    out->elem_n = n[KK_ELEM]; out->node_n = n[KK_NODE]; out->geo_n = n[KK_GEO]; out->fin_n = n[KK_FIN];
    out->heat_elem_ms = ms[KK_HEAT_ELEM]; out->heat_node_ms = ms[KK_HEAT_NODE];
    out->heat_elem_n = n[KK_HEAT_ELEM]; out->heat_node_n = n[KK_HEAT_NODE];
    return SNH_OK;
}

/* plain 16-byte copy, 4 loads in flight per lane, one pass over the buffer */
constexpr int COPY_BLOCK = 256, COPY_UNROLL = 4;
typedef double copy_vec __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(COPY_BLOCK) k_copy(const copy_vec* __restrict__ src, copy_vec* __restrict__ dst,
                                                     size_t n) {
    const size_t base = (size_t)blockIdx.x * (COPY_BLOCK * COPY_UNROLL) + threadIdx.x;
    copy_vec r[COPY_UNROLL];
#pragma unroll
    for (int u = 0; u < COPY_UNROLL; u++) {
        size_t i = base + (size_t)u * COPY_BLOCK;
        if (i < n) r[u] = __builtin_nontemporal_load(&src[i]);
    }
#pragma unroll
    for (int u = 0; u < COPY_UNROLL; u++) {
        size_t i = base + (size_t)u * COPY_BLOCK;
        if (i < n) __builtin_nontemporal_store(r[u], &dst[i]);
    }
}

extern "C" int snh_measure_copy_bandwidth(size_t nbytes, int iters, double* gbps) {
    if (!gbps || nbytes < 1024 || iters < 1) return set_err(SNH_EINVAL, "bad arguments");
    size_t n = nbytes / sizeof(copy_vec);
    copy_vec *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void**)&a, n * sizeof(copy_vec)));
    if (hipMalloc((void**)&b, n * sizeof(copy_vec)) != hipSuccess) {
        (void)hipFree(a);
        return set_err(SNH_ENOMEM, "copy probe allocation failed");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemset(a, 0, n * sizeof(copy_vec));
    const unsigned blocks = (unsigned)((n + COPY_BLOCK * COPY_UNROLL - 1) / (COPY_BLOCK * COPY_UNROLL));
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(COPY_BLOCK), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int it = 0; it < iters; it++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(COPY_BLOCK), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return set_err(SNH_EHIP, "copy probe failed: %s", hipGetErrorString(err));
    *gbps = 2.0 * (double)(n * sizeof(copy_vec)) * iters / (ms * 1e-3) / 1e9;
    return SNH_OK;
}

extern "C" int snh_set_elem_variant(snh_context* c, int32_t v) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (v < 0 || v >= SNH_ELEM_VARIANTS) return set_err(SNH_EINVAL, "unknown element variant");
    c->elem_variant = v;
    return SNH_OK;
}

extern "C" int snh_set_node_variant(snh_context* c, int32_t v) {
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (v != SNH_NODE_PER_NODE && v != SNH_NODE_TILED) return set_err(SNH_EINVAL, "unknown node variant");
    c->node_variant = v;
    return SNH_OK;
}

static int tile_blocks(const snh_context* c, int kc) {
    return ((c->I + NT_I - 1) / NT_I) * ((c->J + NT_J - 1) / NT_J) * ((c->K + kc - 1) / kc);
}

/* the FP64 node pass: the per-node gather and the tiled pass with 4, 8 and
   16 node planes per workgroup, each timed on the context's own state with
   its velocity, coordinate and partial-maximum outputs in scratch; keeps the
   fastest (all give the same bits) */
static int autotune_node(snh_context* c, snh_tune_report& rep) {
    static const int kcs[SNH_NODE_TUNES - 1] = {4, 8, 16};
    const size_t N = c->N;
    int maxb = c->n_node_blocks;
    for (int kc : kcs) maxb = std::max(maxb, tile_blocks(c, kc));
    double *sv = nullptr, *sx = nullptr, *spv = nullptr;
    bool ok = hipMalloc((void**)&sv, 3 * N * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&sx, 3 * N * sizeof(double)) == hipSuccess &&
              hipMalloc((void**)&spv, (size_t)maxb * sizeof(double)) == hipSuccess;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    int rc = ok ? SNH_OK : set_err(SNH_ENOMEM, "autotune scratch allocation failed");
    const int saved_v = c->node_variant, saved_kc = c->node_kc, saved_tb = c->n_tile_blocks;
    for (int o = 0; o < SNH_NODE_TUNES && rc == SNH_OK; o++) {
        c->node_variant = o == 0 ? SNH_NODE_PER_NODE : SNH_NODE_TILED;
        c->node_kc = o == 0 ? saved_kc : kcs[o - 1];
        c->n_tile_blocks = tile_blocks(c, c->node_kc);
        StepArgs a = make_args(c);
        a.v_out = sv;
        a.xn = sx;
        a.part_vmax = spv;
        a.force_out = nullptr;
        float best = 1e30f;
        for (int r = 0; r < 4 && rc == SNH_OK; r++) {
            (void)hipEventRecord(e0, c->stream);
            launch_node_fp64(c, a);
            (void)hipEventRecord(e1, c->stream);
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                rc = set_err(SNH_EHIP, "autotune node launch failed");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0 && ms < best) best = ms;   /* the first launch is a warm-up */
        }
        rep.node_ms[o] = best;
    }
    c->node_variant = saved_v;
    c->node_kc = saved_kc;
    c->n_tile_blocks = saved_tb;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    void* p[] = {sv, sx, spv};
    for (void* q : p)
        if (q) (void)hipFree(q);
    if (rc) return rc;
    int best = 0;
    for (int o = 1; o < SNH_NODE_TUNES; o++)
        if (rep.node_ms[o] < rep.node_ms[best]) best = o;
    const int kc = best == 0 ? saved_kc : kcs[best - 1];
    const int tb = tile_blocks(c, kc);
    if (best > 0 && tb > std::max(c->n_node_blocks, saved_tb)) {
        /* more block maxima than the context holds room for */
        double* pv = nullptr;
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMalloc((void**)&pv, (size_t)tb * sizeof(double)));
        (void)hipFree(c->part_vmax);
        c->part_vmax = pv;
    }
    c->node_variant = best == 0 ? SNH_NODE_PER_NODE : SNH_NODE_TILED;
    c->node_kc = kc;
    c->n_tile_blocks = tb;
    rep.node_variant = c->node_variant;
    rep.node_kc = kc;
    return SNH_OK;
}

extern "C" int snh_autotune(snh_context* c, snh_tune_report* out) {
    Range range_("snh_autotune");
    if (!c) return set_err(SNH_EINVAL, "null context");
    if (!c->zero_done) return set_err(SNH_ESTATE, "snh_autotune before snh_timestep_zero");
    snh_tune_report rep;
    memset(&rep, 0, sizeof rep);
    const size_t E = c->E;
    double *st = nullptr, *ps = nullptr, *visc = nullptr;
    HIPCHK(hipMalloc((void**)&st, 60 * E * sizeof(double)));
    if (c->ps && hipMalloc((void**)&ps, 10 * E * sizeof(double)) != hipSuccess) {
        (void)hipFree(st);
        return set_err(SNH_ENOMEM, "autotune scratch allocation failed");
    }
    if (c->visc && hipMalloc((void**)&visc, 10 * E * sizeof(double)) != hipSuccess) {
        (void)hipFree(st);
        (void)hipFree(ps);
        return set_err(SNH_ENOMEM, "autotune scratch allocation failed");
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int saved = c->elem_variant;
    const int reps = 3;
    int rc = SNH_OK;
    for (int v = 0; v < SNH_ELEM_VARIANTS && rc == SNH_OK; v++) {
        c->elem_variant = v;
        StepArgs a = make_args(c);
        a.stress = st;
        a.stress_out = st;
        a.ps = ps;
        a.ps_out = ps;
        a.visc = visc;
        float best = 1e30f;
        for (int r = 0; r < reps + 1 && rc == SNH_OK; r++) {
            /* fresh copy of the state each launch: the kernel updates in place */
            if (hipMemcpyAsync(st, c->stress, 60 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
                (ps && hipMemcpyAsync(ps, c->ps, 10 * E * sizeof(double), hipMemcpyDeviceToDevice, c->stream) != hipSuccess)) {
                rc = set_err(SNH_EHIP, "autotune copy failed");
                break;
            }
            (void)hipEventRecord(e0, c->stream);
            rc = launch_elem_args(c, a);
            (void)hipEventRecord(e1, c->stream);
            if (hipEventSynchronize(e1) != hipSuccess) rc = set_err(SNH_EHIP, "autotune launch failed");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0 && ms < best) best = ms;   /* first launch is a warm-up */
        }
        rep.elem_ms[v] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(st);
    if (ps) (void)hipFree(ps);
    if (visc) (void)hipFree(visc);
    if (rc) {
        c->elem_variant = saved;
        return rc;
    }
    int best = 0;
    for (int v = 1; v < SNH_ELEM_VARIANTS; v++)
        if (rep.elem_ms[v] < rep.elem_ms[best]) best = v;
    c->elem_variant = best;
    if (getenv("SNH_ELEM_VARIANT")) c->elem_variant = saved;
    rep.elem_variant = c->elem_variant;
    rep.node_variant = c->node_variant;
    rep.node_kc = c->node_kc;
    if (c->cfg.precision == SNH_PRECISION_FP64 && !getenv("SNH_NODE_VARIANT") && (rc = autotune_node(c, rep)))
        return rc;
    if (out) *out = rep;
    return SNH_OK;
}

/* ==========================================================================
 * snh_sync_state: the Sync of a state in caller device memory, allocating
 * nothing.  The block partials of the minimum length scale and of max |v|
 * and the step scalars live in one fixed device array (`scratch`), whose
 * device address each GPU resolves once with hipGetSymbolAddress.
 * ======================================================================== */
#define SYNC_MAX_ELEMENTS (1 << 22)
#define SYNC_MAX_NODES (1 << 22)
#define SYNC_GEO_PARTS (SYNC_MAX_ELEMENTS / GEO_BLOCK)    /* k_geo block minima */
#define SYNC_NODE_PARTS (SYNC_MAX_NODES / NODE_BLOCK)     /* k_vmax_walls block maxima */
#define SYNC_SCAL_DOUBLES 16                              /* DevScalars, rounded up */
static_assert(sizeof(DevScalars) <= SYNC_SCAL_DOUBLES * sizeof(double), "scalars do not fit the scratch");
static __device__ __align__(256) double scratch[SYNC_GEO_PARTS + SYNC_NODE_PARTS + SYNC_SCAL_DOUBLES];

#define SYNC_MAX_DEVICES 64
static double* g_scratch_ptr[SYNC_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* scratch's address on the current device, resolved on the first call there */
static int sync_scratch(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= SYNC_MAX_DEVICES) return set_err(SNH_EUNSUPPORTED, "device %d beyond %d", dev, SYNC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double* p = nullptr;
        hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(SNH_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SNH_OK;
}

/* the scalars as snh_context_create sets them: dt = the configured timeStep */
__global__ void k_sync_init(DevScalars* s, double dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->dt = dt;
        s->dt_rot = dt;
        s->speedOfSound = 0.0;
        s->minLen = 0.0;
        s->vmax = 0.0;
        s->thermal_err = 0;
        s->nonfinite = 0;
        s->topoGradMax = 0.0;
    }
}

/* block maxima of |v| over the nodes on the walls of the velocity
   conditions (bit w of `walls` for SNH_WALL_w), as k_vmax over the
   context's BC codes */
__global__ void __launch_bounds__(NODE_BLOCK) k_vmax_walls(StepArgs a, uint32_t walls) {
    __shared__ double sh[NODE_BLOCK / 64];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    double vmag = -INFINITY;
    if (n < a.N) {
        int i, j, k;
        node_ijk(a, n, i, j, k);
        const bool on = ((walls >> SNH_WALL_ALL) & 1u) || (((walls >> SNH_WALL_FRONT) & 1u) && a.K > 1 && k == 0) ||
                        (((walls >> SNH_WALL_BACK) & 1u) && a.K > 1 && k == a.K - 1) ||
                        (((walls >> SNH_WALL_LEFT) & 1u) && a.I > 1 && i == 0) ||
                        (((walls >> SNH_WALL_RIGHT) & 1u) && a.I > 1 && i == a.I - 1) ||
                        (((walls >> SNH_WALL_TOP) & 1u) && a.J > 1 && j == a.J - 1) ||
                        (((walls >> SNH_WALL_BOTTOM) & 1u) && a.J > 1 && j == 0);
        if (on) {
            const size_t N = (size_t)a.N;
            const double v0 = a.v[n], v1 = a.v[N + n], v2 = a.v[2 * N + n];
            vmag = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
        }
    }
    const double bm = block_max(vmag, sh);
    if (threadIdx.x == 0) a.part_vmax[blockIdx.x] = bm;
}

__global__ void k_sync_out(const DevScalars* s, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        out[0] = s->dt;
        out[1] = s->speedOfSound;
        out[2] = s->minLen;
        out[3] = s->vmax;
    }
}

extern "C" int snh_sync_state(const snh_config* cfg, const double* coord, const double* velocity, double* out,
                              void* stream) {
    Range range_("snh_sync_state");
    if (!cfg) return set_err(SNH_EINVAL, "snh_sync_state: null config");
    if (cfg->meshSizeI < 2 || cfg->meshSizeJ < 2 || cfg->meshSizeK < 2)
        return set_err(SNH_EINVAL, "mesh needs at least 2 nodes per axis");
    const size_t I = cfg->meshSizeI, J = cfg->meshSizeJ, K = cfg->meshSizeK;
    const size_t N = I * J * K, E = (I - 1) * (J - 1) * (K - 1);
    if (N > SYNC_MAX_NODES || E > SYNC_MAX_ELEMENTS)
        return set_err(SNH_EINVAL, "snh_sync_state: %zu nodes / %zu elements beyond its %d / %d", N, E,
                       SYNC_MAX_NODES, SYNC_MAX_ELEMENTS);
    if (cfg->dtType != SNH_DT_DYNAMIC && cfg->dtType != SNH_DT_CONSTANT && cfg->dtType != SNH_DT_WAVE)
        return set_err(SNH_EINVAL, "bad dtType");
    if (cfg->dtType == SNH_DT_WAVE && (cfg->nmaterials < 1 || !cfg->materials))
        return set_err(SNH_EINVAL, "the wave dt type needs the materials");
    if (cfg->nbcs && !cfg->bcs) return set_err(SNH_EINVAL, "nbcs > 0 with null bcs");
    if (!coord || !velocity || !out) return set_err(SNH_EINVAL, "snh_sync_state: null device pointer");
    uint32_t walls = 0;
    for (uint32_t b = 0; b < cfg->nbcs; b++) {
        if (cfg->bcs[b].wall < SNH_WALL_FRONT || cfg->bcs[b].wall > SNH_WALL_ALL)
            return set_err(SNH_EINVAL, "bad wall %d", cfg->bcs[b].wall);
        walls |= 1u << cfg->bcs[b].wall;
    }
    double* sp = nullptr;
    int rc = sync_scratch(&sp);
    if (rc) return rc;
    StepArgs a;
    memset(&a, 0, sizeof a);
    a.xc = coord;
    a.v = velocity;
    a.part_min = sp;
    a.part_vmax = sp + SYNC_GEO_PARTS;
    a.scal = (DevScalars*)(sp + SYNC_GEO_PARTS + SYNC_NODE_PARTS);
    a.I = (int)I;
    a.J = (int)J;
    a.K = (int)K;
    a.EI = (int)I - 1;
    a.EJ = (int)J - 1;
    a.EK = (int)K - 1;
    a.N = (int)N;
    a.E = (int)E;
    a.dtType = cfg->dtType;
    a.strain_inert = cfg->strain_inert;
    a.c_wave = 0.0;
    for (uint32_t q = 0; q < cfg->nmaterials && cfg->materials; q++) a.c_wave = wave_speed_max(a.c_wave, cfg->materials[q]);
    set_divisors(a);
    hipStream_t st = (hipStream_t)stream;
    const int ngeo = (int)((E + GEO_BLOCK - 1) / GEO_BLOCK), nnode = (int)((N + NODE_BLOCK - 1) / NODE_BLOCK);
    hipLaunchKernelGGL(k_sync_init, dim3(1), dim3(64), 0, st, a.scal, cfg->timeStep);
    hipLaunchKernelGGL(k_vmax_walls, dim3(nnode), dim3(NODE_BLOCK), 0, st, a, walls);
    hipLaunchKernelGGL(k_geo, dim3(ngeo), dim3(GEO_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(FIN_BLOCK), 0, st, a, ngeo, nnode);
    hipLaunchKernelGGL(k_sync_out, dim3(1), dim3(64), 0, st, (const DevScalars*)a.scal, out);
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

#include "snh_adjoint_host.h"

/* ==========================================================================
 * ensembles: one HIP graph steps every member, members on parallel branches
 * ======================================================================== */
/* a group of members stepped by batched launches on one stream: the tables
   of BatchGrid for the element, node and geometry passes */
struct EnsGroup {
    std::vector<int> idx;                      /* members of the ensemble       */
    hipStream_t stream = nullptr;
    hipEvent_t join = nullptr;
    StepArgs base[2];                          /* first member's StepArgs, parity q */
    BatchMember* dmem[2] = {nullptr, nullptr}; /* members relative to it, parity q  */
    int* dfirst[3] = {nullptr, nullptr, nullptr};
    int* dcount[3] = {nullptr, nullptr, nullptr};
    uint16_t* dmember[3] = {nullptr, nullptr, nullptr};
    int grid[3] = {0, 0, 0};
};

struct snh_ensemble {
    std::vector<snh_context*> m;
    std::vector<int> cur0;      /* members' coordinate-buffer index at capture */
    int graph_steps;
    hipStream_t origin;
    hipEvent_t fork;
    std::vector<hipEvent_t> join;
    /* [0]: buffers as at capture, [1]: every member's buffers swapped */
    hipGraph_t graph[2];
    hipGraphExec_t exec[2];
    /* batched launches (members of one precision and law): the members in
       groups, each group's passes batched on a stream of its own */
    bool batched = false;
    bool tw = false;            /* element pass layout of the batched launches:
                                   tet per lane (tet per wave with
                                   SNH_ENSEMBLE_TW=1) */
    std::vector<EnsGroup> groups;
};

enum { EB_ELEM = 0, EB_NODE = 1, EB_GEO = 2 };

static void ens_free(snh_ensemble* e) {
    if (!e) return;
    for (int p = 0; p < 2; p++) {
        if (e->exec[p]) (void)hipGraphExecDestroy(e->exec[p]);
        if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
    }
    if (e->fork) (void)hipEventDestroy(e->fork);
    for (hipEvent_t j : e->join)
        if (j) (void)hipEventDestroy(j);
    if (e->origin) (void)hipStreamDestroy(e->origin);
    for (EnsGroup& g : e->groups) {
        for (int q = 0; q < 2; q++)
            if (g.dmem[q]) (void)hipFree(g.dmem[q]);
        for (int k = 0; k < 3; k++) {
            if (g.dfirst[k]) (void)hipFree(g.dfirst[k]);
            if (g.dcount[k]) (void)hipFree(g.dcount[k]);
            if (g.dmember[k]) (void)hipFree(g.dmember[k]);
        }
        if (g.join) (void)hipEventDestroy(g.join);
        if (g.stream) (void)hipStreamDestroy(g.stream);
    }
    delete e;
}

/* Members that one batched launch per pass can step: one precision, one
   single law, one material layout (with or without an element material
   map), the per-node FP64 node pass, no dump diagnostics or tet densities.
   Batched launches are taken when the members average fewer than 16384
   elements: there the launches set the time (256 models of ~900 elements:
   8x the per-member branches), while with larger members the per-member
   branches, each in its autotuned layout and overlapping on the hardware
   queues, are as fast or faster (64 models of 53k elements on average: FP64
   1.03x, mixed 1.10x the batched launches).  SNH_ENSEMBLE_BATCH=1 in the
   environment batches whatever the sizes, =0 never. */
static bool ens_batchable(const std::vector<snh_context*>& m) {
    const char* env = getenv("SNH_ENSEMBLE_BATCH");
    if (env && env[0] == '0') return false;
    const bool force = env && env[0] == '1';
    if (m.size() > 65535) return false;
    size_t total = 0;
    for (const snh_context* c : m) total += c->E;
    if (!force && total >= (size_t)16384 * m.size()) return false;
    const snh_context* c0 = m[0];
    if (c0->rset != RS_ELASTIC && c0->rset != RS_PLASTIC && c0->rset != RS_MAXWELL) return false;
    for (const snh_context* c : m) {
        if (c->cfg.precision != c0->cfg.precision || c->rset != c0->rset ||
            (c->mat != nullptr) != (c0->mat != nullptr))
            return false;
        if (c->diag_on || c->beta_dens || c->thermal || c->node_variant != SNH_NODE_PER_NODE) return false;
        if (c->cfg.precision == SNH_PRECISION_MIXED && c->mat) return false;
    }
    return true;
}

/* host -> device copy ordered on the ensemble's stream, complete on return */
static hipError_t ens_upload(snh_ensemble* e, void* dst, const void* src, size_t bytes) {
    hipError_t r = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->origin);
    return r != hipSuccess ? r : hipStreamSynchronize(e->origin);
}

/* the members in groups of about equal element counts (largest first, each
   to the lightest group): each group's batched passes run on a stream of
   its own, so one group's element pass may overlap another's node pass.
   SNH_ENSEMBLE_GROUPS sets the count (default 1: on the 64-model ensemble
   one group is the fastest; on hundreds of tiny models four are). */
static int ens_make_groups(snh_ensemble* e) {
    const int n = (int)e->m.size();
    int ng = 1;
    if (const char* env = getenv("SNH_ENSEMBLE_GROUPS")) ng = atoi(env);
    if (ng < 1) ng = 1;
    if (ng > n) ng = n;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return e->m[a]->E > e->m[b]->E; });
    e->groups.resize(ng);
    std::vector<size_t> load(ng, 0);
    for (int i : order) {
        int g = 0;
        for (int k = 1; k < ng; k++)
            if (load[k] < load[g]) g = k;
        e->groups[g].idx.push_back(i);
        load[g] += e->m[i]->E;
    }
    for (EnsGroup& g : e->groups) {
        std::sort(g.idx.begin(), g.idx.end());
        HIPCHK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g.join, hipEventDisableTiming));
    }
    return SNH_OK;
}

/* device tables of a group's batched launches; its members' StepArgs for
   both coordinate-buffer parities, relative to its first member's */
static int ens_batch_tables(snh_ensemble* e, EnsGroup& grp) {
    const int n = (int)grp.idx.size();
    std::vector<int> cnt[3], first[3];
    for (int k = 0; k < 3; k++) {
        cnt[k].resize(n);
        first[k].resize(n);
    }
    for (int i = 0; i < n; i++) {
        const snh_context* c = e->m[grp.idx[i]];
        cnt[EB_ELEM][i] = (int)(e->tw ? (c->E + 63) / 64 : (c->E + EPW - 1) / EPW);
        cnt[EB_NODE][i] = c->n_node_blocks;
        cnt[EB_GEO][i] = c->n_geo_blocks;
    }
    for (int k = 0; k < 3; k++) {
        int64_t at = 0;
        for (int i = 0; i < n; i++) {
            first[k][i] = (int)at;
            at += ((int64_t)cnt[k][i] + 7) / 8 * 8;
        }
        if (at > INT32_MAX) return set_err(SNH_EUNSUPPORTED, "ensemble too large for one batched launch");
        grp.grid[k] = (int)at;
        std::vector<uint16_t> mem((size_t)at);
        for (int i = 0; i < n; i++)
            for (int64_t b = first[k][i]; b < first[k][i] + ((int64_t)cnt[k][i] + 7) / 8 * 8; b++)
                mem[(size_t)b] = (uint16_t)i;
        HIPCHK(hipMalloc((void**)&grp.dfirst[k], n * sizeof(int)));
        HIPCHK(hipMalloc((void**)&grp.dcount[k], n * sizeof(int)));
        HIPCHK(hipMalloc((void**)&grp.dmember[k], mem.size() * sizeof(uint16_t)));
        HIPCHK(ens_upload(e, grp.dfirst[k], first[k].data(), n * sizeof(int)));
        HIPCHK(ens_upload(e, grp.dcount[k], cnt[k].data(), n * sizeof(int)));
        HIPCHK(ens_upload(e, grp.dmember[k], mem.data(), mem.size() * sizeof(uint16_t)));
    }
    for (int q = 0; q < 2; q++) {
        std::vector<StepArgs> args(n);
        for (int i = 0; i < n; i++) {
            snh_context* c = e->m[grp.idx[i]];
            const int cur = c->cur;
            c->cur = e->cur0[grp.idx[i]] ^ q;
            args[i] = make_args(c);
            c->cur = cur;
        }
        const StepArgs& b = args[0];
        grp.base[q] = b;
        std::vector<BatchMember> mem(n);
        for (int i = 0; i < n; i++) {
            const StepArgs& a = args[i];
            BatchMember& d = mem[i];
#define SNH_DELTA(f) d.f = (int64_t)((intptr_t)a.f - (intptr_t)b.f)
            SNH_DELTA(xc);
            SNH_DELTA(xp);
            SNH_DELTA(xn);
            SNH_DELTA(v);
            SNH_DELTA(v_out);
            SNH_DELTA(stress);
            SNH_DELTA(stress_out);
            SNH_DELTA(ps);
            SNH_DELTA(ps_out);
            SNH_DELTA(visc);
            SNH_DELTA(mat);
            SNH_DELTA(refvisc);
            SNH_DELTA(cohoff);
            SNH_DELTA(bc);
            SNH_DELTA(tape_f);
            SNH_DELTA(tape_m);
            SNH_DELTA(tape_code);
            SNH_DELTA(damp_sched);
            SNH_DELTA(diag);
            SNH_DELTA(force_out);
            SNH_DELTA(temp);
            SNH_DELTA(temp0);
            SNH_DELTA(tpre);
            SNH_DELTA(hflux);
            SNH_DELTA(tstress);
            SNH_DELTA(dens);
            SNH_DELTA(tbc);
            SNH_DELTA(tbcv);
            SNH_DELTA(mats);
            SNH_DELTA(bcs);
            SNH_DELTA(part_min);
            SNH_DELTA(part_vmax);
            SNH_DELTA(scal);
#undef SNH_DELTA
            d.I = a.I;
            d.J = a.J;
            d.K = a.K;
            d.EI = a.EI;
            d.EJ = a.EJ;
            d.EK = a.EK;
            d.N = a.N;
            d.E = a.E;
            d.dtType = a.dtType;
            d.node_band = a.node_band;
            d.node_kc = a.node_kc;
            d.gravity = a.gravity;
            d.demf = a.demf;
            d.strain_inert = a.strain_inert;
            d.c_wave = a.c_wave;
            d.dI = a.dI;
            d.dJ = a.dJ;
            d.dEI = a.dEI;
            d.dEJ = a.dEJ;
            d.dKB = a.dKB;
            d.dPer = a.dPer;
        }
        HIPCHK(hipMalloc((void**)&grp.dmem[q], n * sizeof(BatchMember)));
        HIPCHK(ens_upload(e, grp.dmem[q], mem.data(), n * sizeof(BatchMember)));
    }
    return SNH_OK;
}

/* one step of a group's members, batched, on stream s, entering with
   buffer parity q: the element and node passes read the coordinates of
   parity q, the node pass writes the other buffer, and the geometry pass
   and Sync see the buffers swapped (as snh_step's launches do after the
   node pass flips the context's index) */
static int ens_batch_step(const snh_ensemble* e, const EnsGroup& grp, hipStream_t s, int q) {
    const snh_context* c0 = e->m[grp.idx[0]];
    const bool mm = c0->mat != nullptr;
    BatchGrid g[3];
    for (int k = 0; k < 3; k++) {
        const int p = k == EB_GEO ? q ^ 1 : q;
        g[k] = BatchGrid{grp.base[p], grp.dmem[p], grp.dfirst[k], grp.dcount[k], grp.dmember[k]};
    }
    const dim3 ge(grp.grid[EB_ELEM]), gn(grp.grid[EB_NODE]), gg(grp.grid[EB_GEO]);
    const dim3 be(e->tw ? 640 : 64);
    const bool tw = e->tw;
    if (c0->cfg.precision == SNH_PRECISION_MIXED) {
#define EB_MX_LAUNCH(R)                                                                          \
    do {                                                                                         \
        if (tw) hipLaunchKernelGGL((k_elem_mx_batch<R, true>), ge, be, 0, s, g[EB_ELEM]);        \
        else hipLaunchKernelGGL((k_elem_mx_batch<R, false>), ge, be, 0, s, g[EB_ELEM]);          \
    } while (0)
        switch (c0->rset) {
        case RS_ELASTIC: EB_MX_LAUNCH(RS_ELASTIC); break;
        case RS_PLASTIC: EB_MX_LAUNCH(RS_PLASTIC); break;
        default: EB_MX_LAUNCH(RS_MAXWELL); break;
        }
#undef EB_MX_LAUNCH
        hipLaunchKernelGGL(k_node_mx_batch, gn, dim3(NODE_BLOCK), 0, s, g[EB_NODE]);
    } else {
#define EB_ELEM_LAUNCH(R)                                                                        \
    do {                                                                                         \
        if (mm) {                                                                                \
            if (tw) hipLaunchKernelGGL((k_elem_batch<R, true, true>), ge, be, 0, s, g[EB_ELEM]); \
            else hipLaunchKernelGGL((k_elem_batch<R, false, true>), ge, be, 0, s, g[EB_ELEM]);   \
        } else {                                                                                 \
            if (tw) hipLaunchKernelGGL((k_elem_batch<R, true, false>), ge, be, 0, s, g[EB_ELEM]); \
            else hipLaunchKernelGGL((k_elem_batch<R, false, false>), ge, be, 0, s, g[EB_ELEM]);  \
        }                                                                                        \
    } while (0)
        switch (c0->rset) {
        case RS_ELASTIC: EB_ELEM_LAUNCH(RS_ELASTIC); break;
        case RS_PLASTIC: EB_ELEM_LAUNCH(RS_PLASTIC); break;
        default: EB_ELEM_LAUNCH(RS_MAXWELL); break;
        }
#undef EB_ELEM_LAUNCH
        if (mm) hipLaunchKernelGGL(k_node_batch<true>, gn, dim3(NODE_BLOCK), 0, s, g[EB_NODE]);
        else hipLaunchKernelGGL(k_node_batch<false>, gn, dim3(NODE_BLOCK), 0, s, g[EB_NODE]);
    }
    hipLaunchKernelGGL(k_geo_batch, gg, dim3(GEO_BLOCK), 0, s, g[EB_GEO]);
    hipLaunchKernelGGL(k_fin_batch, dim3(grp.idx.size()), dim3(FIN_BLOCK), 0, s, grp.base[q ^ 1],
                       (const BatchMember*)grp.dmem[q ^ 1], (const int*)grp.dcount[EB_GEO],
                       (const int*)grp.dcount[EB_NODE]);
    HIPCHK(hipGetLastError());
    return SNH_OK;
}

extern "C" int snh_ensemble_create(snh_context* const* members, int32_t n, int32_t graph_steps,
                                   snh_ensemble** out) {
    if (!members || !out || n < 1) return set_err(SNH_EINVAL, "bad ensemble arguments");
    if (graph_steps < 2 || (graph_steps & 1)) return set_err(SNH_EINVAL, "graph_steps must be even and >= 2");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        snh_context* c = members[i];
        if (!c) return set_err(SNH_EINVAL, "null ensemble member %d", i);
        if (!c->zero_done) return set_err(SNH_ESTATE, "ensemble member %d before snh_timestep_zero", i);
        if (c->sched_n > 0 || c->profiling || c->adj)
            return set_err(SNH_ESTATE, "ensemble member %d has a schedule, profiling or an adjoint", i);
        if (c->thermal) return set_err(SNH_EUNSUPPORTED, "ensemble member %d runs the temperature plugin", i);
        if (c->cfg.topo_kappa > 0)
            return set_err(SNH_EUNSUPPORTED, "ensemble member %d diffuses its topography (topo_kappa > 0)", i);
        for (int j = 0; j < i; j++)
            if (members[j] == c || members[j]->stream == c->stream)
                return set_err(SNH_EINVAL, "ensemble members need distinct contexts and streams");
    }
    snh_ensemble* e = new (std::nothrow) snh_ensemble();
    if (!e) return set_err(SNH_ENOMEM, "out of host memory");
    e->m.assign(members, members + n);
    e->graph_steps = graph_steps;
    e->origin = nullptr;
    e->fork = nullptr;
    e->graph[0] = e->graph[1] = nullptr;
    e->exec[0] = e->exec[1] = nullptr;
    e->join.assign(n, nullptr);
    e->cur0.resize(n);
    for (int i = 0; i < n; i++) e->cur0[i] = members[i]->cur;
    int rc = SNH_OK;
    do {
        if (hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
            rc = set_err(SNH_EHIP, "ensemble stream/event creation failed");
            break;
        }
        for (int i = 0; i < n && rc == SNH_OK; i++)
            if (hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) != hipSuccess)
                rc = set_err(SNH_EHIP, "ensemble event creation failed");
        if (rc) break;
        /* members' pending work first: the graph then starts from a settled state */
        for (snh_context* c : e->m)
            if (hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(SNH_EHIP, "member sync failed");
        if (rc) break;
        e->batched = ens_batchable(e->m);
        if (const char* env = getenv("SNH_ENSEMBLE_TW")) e->tw = env[0] != '0';
        if (e->batched && (rc = ens_make_groups(e))) break;
        for (size_t k = 0; e->batched && k < e->groups.size() && rc == SNH_OK; k++)
            rc = ens_batch_tables(e, e->groups[k]);
        if (rc) break;
        /* capture records the launches without running them; the host-side
           buffer index is set for each of the two parities and restored */
        for (int par = 0; par < 2 && rc == SNH_OK; par++) {
            for (int i = 0; i < n; i++) e->m[i]->cur = e->cur0[i] ^ par;
            if (hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = set_err(SNH_EHIP, "hipStreamBeginCapture failed");
                break;
            }
            int crc = SNH_OK;
            if (hipEventRecord(e->fork, e->origin) != hipSuccess) crc = SNH_EHIP;
            std::vector<uint32_t> t0(n);
            for (int i = 0; i < n; i++) t0[i] = e->m[i]->timeStep;
            /* batched: each group's steps on its own stream, forked from and
               joined to the origin */
            for (size_t k = 0; e->batched && k < e->groups.size() && crc == SNH_OK; k++) {
                const EnsGroup& grp = e->groups[k];
                if (hipStreamWaitEvent(grp.stream, e->fork, 0) != hipSuccess) crc = SNH_EHIP;
                for (int s = 0; s < graph_steps && crc == SNH_OK; s++)
                    crc = ens_batch_step(e, grp, grp.stream, par ^ (s & 1));
                if (crc == SNH_OK && hipEventRecord(grp.join, grp.stream) != hipSuccess) crc = SNH_EHIP;
                if (crc == SNH_OK && hipStreamWaitEvent(e->origin, grp.join, 0) != hipSuccess) crc = SNH_EHIP;
            }
            for (int i = 0; !e->batched && i < n && crc == SNH_OK; i++) {
                snh_context* c = e->m[i];
                t0[i] = c->timeStep;
                if (hipStreamWaitEvent(c->stream, e->fork, 0) != hipSuccess) crc = SNH_EHIP;
                if (crc == SNH_OK) crc = snh_step(c, graph_steps);
                if (crc == SNH_OK && hipEventRecord(e->join[i], c->stream) != hipSuccess) crc = SNH_EHIP;
                if (crc == SNH_OK && hipStreamWaitEvent(e->origin, e->join[i], 0) != hipSuccess) crc = SNH_EHIP;
            }
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(e->origin, &g);
            for (int i = 0; i < n; i++) e->m[i]->timeStep = t0[i];
            if (crc != SNH_OK || ce != hipSuccess || !g) {
                if (g) (void)hipGraphDestroy(g);
                rc = set_err(SNH_EHIP, "ensemble capture failed");
                break;
            }
            e->graph[par] = g;
            if (hipGraphInstantiate(&e->exec[par], g, nullptr, nullptr, 0) != hipSuccess)
                rc = set_err(SNH_EHIP, "hipGraphInstantiate failed");
        }
        for (int i = 0; i < n; i++) e->m[i]->cur = e->cur0[i];
    } while (0);
    if (rc) {
        ens_free(e);
        return rc;
    }
    *out = e;
    return SNH_OK;
}

extern "C" int snh_ensemble_step(snh_ensemble* e, int32_t nsteps) {
    if (!e) return set_err(SNH_EINVAL, "null ensemble");
    if (nsteps < 0) return set_err(SNH_EINVAL, "negative step count");
    /* every member's buffers as at capture, or every member's swapped:
       otherwise (members stepped on their own meanwhile) plain steps */
    const int par = e->m[0]->cur ^ e->cur0[0];
    bool same = true;
    for (size_t i = 0; i < e->m.size(); i++) same = same && ((e->m[i]->cur ^ e->cur0[i]) == par);
    const int reps = same ? nsteps / e->graph_steps : 0;
    const int rest = nsteps - reps * e->graph_steps;
    if (reps > 0)
        for (size_t i = 0; i < e->m.size(); i++) {   /* graph after the members' earlier work */
            HIPCHK(hipEventRecord(e->join[i], e->m[i]->stream));
            HIPCHK(hipStreamWaitEvent(e->origin, e->join[i], 0));
        }
    for (int r = 0; r < reps; r++) {
        HIPCHK(hipGraphLaunch(e->exec[par], e->origin));
        for (snh_context* c : e->m) {
            c->timeStep += e->graph_steps;
            c->rotation_pending = true;
        }
    }
    if (rest) {
        /* the members' own streams wait for the graph, then run the rest */
        HIPCHK(hipEventRecord(e->fork, e->origin));
        for (size_t i = 0; i < e->m.size(); i++) {
            snh_context* c = e->m[i];
            HIPCHK(hipStreamWaitEvent(c->stream, e->fork, 0));
            int rc = snh_step(c, rest);
            if (rc) return rc;
            HIPCHK(hipEventRecord(e->join[i], c->stream));
            HIPCHK(hipStreamWaitEvent(e->origin, e->join[i], 0));
        }
    } else {
        /* later work on a member's own stream follows the graph */
        HIPCHK(hipEventRecord(e->fork, e->origin));
        for (snh_context* c : e->m) HIPCHK(hipStreamWaitEvent(c->stream, e->fork, 0));
    }
    return SNH_OK;
}

extern "C" int snh_ensemble_batched(const snh_ensemble* e) {
    if (!e) return set_err(SNH_EINVAL, "null ensemble");
    return e->batched ? 1 : 0;
}

extern "C" int snh_ensemble_synchronize(snh_ensemble* e) {
    if (!e) return set_err(SNH_EINVAL, "null ensemble");
    HIPCHK(hipStreamSynchronize(e->origin));
    for (snh_context* c : e->m) HIPCHK(hipStreamSynchronize(c->stream));
    return SNH_OK;
}

extern "C" int snh_ensemble_destroy(snh_ensemble* e) {
    if (!e) return SNH_OK;
    (void)hipStreamSynchronize(e->origin);
    ens_free(e);
    return SNH_OK;
}
