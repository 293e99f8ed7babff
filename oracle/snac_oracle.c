/*
 * snac_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for SNAC's explicit
 * time step.  Restates, in plain C and in the same arithmetic order, the
 * reference functions cited at each definition (paths relative to the
 * geodynamics/snac tree).  Only tests/, the smoke check and bench.py's CPU
 * baseline use it; the HIP product never does.
 *
 * How each part is pinned (tests/test_oracle_*.py, tests/test_tangent_model.py):
 *   connectivity, tet tables      exact tables + invariants (5 tets per node per
 *                                 element, face opposite the node, 8E incidences)
 *   tet geometry, minLengthScale  cube closed forms (h^3/6, h^3/3, h^2/2,
 *                                 sqrt(3)/2 h^2, h/(3 sqrt 3)); reference fixtures
 *                                 testTetrahedra-volume (166667),
 *                                 testStress-stress (minLengthScale 19.245)
 *   strain rate / rotation        affine velocity fields (exact sym/skew L),
 *                                 rigid motions (zero strain rate)
 *   force                         reference fixtures testUpdateNode-force
 *                                 (-3.31088e9 ... -2.6487e10) and -balance
 *                                 (accumulation order); force balance without
 *                                 gravity
 *   masses                        physical mass: fixture testUpdateNode-mass
 *                                 (42.1875 ...); dynamic inertial mass
 *                                 1/4 (lambda+2mu)/c^2 sum V and wave mass
 *                                 1/4 rho sum V: closed forms on box meshes of
 *                                 two materials (sum V = abc per incident element);
 *                                 constant-dt inertial mass: through the
 *                                 testUpdateNode-velocity fixture, which runs
 *                                 with dtType constant
 *   node update, BCs              reference fixtures testUpdateNode-velocity
 *                                 (vx = 0.00834447 after 10 iterations) and
 *                                 -coord (0.550852, 1.1)
 *   dt (Sync)                     test-basic timeStep.0 (1.80422e+07)
 *   element invariant, density    hydrostatic fixture (below); tet density
 *                                 rho0 (1 - alpha (T - T_ref) + beta sigma_m)
 *                                 under a uniform stress, and the gravity force
 *                                 it gives an interior node
 *   hydrostatic IC                the reference's float32 stress dumps of
 *                                 basic-hydroStatic.xml, bit for bit; with beta,
 *                                 second-order convergence to the exponential
 *                                 profile (1 - exp(-beta rho0 g z)) / beta
 *   topography diffusion          parabolic top surface at rest: the
 *                                 five-point stencil's node-order recurrence
 *                                 dh_i = kappa dt (2a + dh_(i-1)/h^2) along x
 *                                 and z, the switch-on node, the edge copies
 *                                 and the running maximum of |grad y|; on
 *                                 meshes two nodes wide in x (z) the
 *                                 three-point recurrence along z (x), the
 *                                 second row starting from the edge node's
 *                                 copy of an old height, and no diffusion
 *                                 two nodes wide in both
 *   elastic                       uniaxial strain closed form
 *   plastic_BI eigen-solver       numpy.linalg.eigh on random symmetric tensors,
 *                                 int-abs truncation cases
 *   plastic_BI return mapping     oedometer benchmark (Snac/tests/oedometer):
 *                                 elastic path exact, post-yield path within
 *                                 3e-5 of analytic_soln.py's corner solution,
 *                                 plastic strain growing as the measure of the
 *                                 flow direction times the analytic Epx; shear
 *                                 return = flow-rule projection with the
 *                                 multiplier fixed by f_s = 0; tensile return
 *                                 onto the cut-off (c/tan(phi) or ten_off)
 *   softening (soften)            per-tet plastic strain increments against
 *                                 numpy.interp of the tables, on, between and
 *                                 beyond the nodes (defaults 30 MPa, 30 deg,
 *                                 0 deg; carry-over between tets)
 *   plastic-strain increment      recovered from the stress change through the
 *                                 elastic compliance (shear and tension)
 *   maxwell                       constant-strain relaxation ((1-a)/(1+a))^n
 *   viscosity (viscosity_of)      power law: steady stress scales as
 *                                 rate^(1/n); Arrhenius: relaxation rate at a
 *                                 uniform temperature
 *   viscoplastic_BI               Maxwell relaxation below yield; return onto
 *                                 the surface of the cohesion at the updated
 *                                 plastic strain (hardening term), where
 *                                 plastic_BI's is at the previous one;
 *                                 diagonal-only eigen-solve on yield (lower
 *                                 triangle zero)
 *   temperature (energy EPs)      the reference's float32 temperature dumps of
 *                                 basic-temp.xml, bit for bit; uniform and
 *                                 linear fields (zero flux / exact gradient,
 *                                 steady interior), insulated-box energy
 *                                 conservation, 1D diffusion decay rate,
 *                                 thermal stress closed form
 *   remesh handoff                identity remesh, running minimum of the
 *                                 length scale, volume jump of a scaled mesh
 *   hypot                         the system libm's (glibc 2.35), as the reference
 *                                 calls it; the correctly rounded value is selectable
 * libm (sin, tan, pow, exp) is the system's glibc, as in the reference's own
 * build.  The device's versions of those four differ from glibc's by an ulp
 * at most, so the GPU paths that call them (softened friction and dilation
 * angles, power-law and Arrhenius viscosity) are held to the FP64 tolerance
 * against this oracle rather than to bit equality.
 */
#define _GNU_SOURCE
#include "snac_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NTET 10

/* ---------------------------------------------------------------------------
 * float literals of the reference, as the doubles it actually uses
 * ------------------------------------------------------------------------- */
#define F_045      ((double)0.45f)     /* Context.c:1251 */
#define F_3EM11    ((double)3.0e-11f)  /* Context.c:1219 */
#define F_1EP27    ((double)1.0e+27f)  /* UpdateElement.c:50 */

/* ---------------------------------------------------------------------------
 * tetrahedra tables (Snac/libSnac/src/TetrahedraTables.c:42-84).  These are
 * the reference's decomposition of the hex into two sets of five tetrahedra;
 * the table must be identical for the stored tet stresses to mean the same.
 * ------------------------------------------------------------------------- */
static const unsigned TetraToNode[NTET][4] = {
    {7, 2, 3, 0}, {0, 5, 4, 7}, {2, 5, 1, 0}, {5, 2, 6, 7}, {2, 5, 0, 7},
    {3, 4, 7, 6}, {4, 3, 0, 1}, {6, 1, 2, 3}, {1, 6, 5, 4}, {1, 3, 6, 4}};
static const unsigned FaceToNode[4][3] = {{0, 1, 2}, {1, 3, 2}, {0, 2, 3}, {3, 1, 0}};
static const unsigned NodeToTetra[8][5] = {
    {3, 5, 7, 8, 9}, {0, 1, 3, 4, 5}, {1, 2, 3, 4, 8}, {1, 5, 6, 8, 9},
    {0, 2, 3, 4, 7}, {0, 5, 6, 7, 9}, {2, 6, 7, 8, 9}, {0, 1, 2, 4, 6}};
static const unsigned NodeToSurface[8][5] = {
    {3, 0, 1, 2, 3}, {1, 0, 0, 0, 3}, {2, 2, 1, 2, 3}, {3, 2, 1, 0, 0},
    {2, 1, 2, 1, 3}, {3, 1, 2, 0, 2}, {3, 0, 2, 1, 1}, {0, 1, 0, 3, 3}};

typedef struct {
    double area;
    double normal[3];
} orc_surface;

typedef struct {
    double old_volume, volume;
    orc_surface surface[4];
    double strainRate[3][3];
    double strain[3][3];
    double stress[3][3];
    double density;
    double avgTemp;
} orc_tet;

typedef struct {
    unsigned material_I;
    double volume, strainRate, stress, hydroPressure;
    orc_tet tetra[NTET];
    /* plugin element extensions (plastic_BI/Element.h:46-47, maxwell/Element.h:45,
       viscoplastic_BI/Element.h:46-48) */
    double plasticStrain[NTET];
    double aps;
    double viscosity[NTET];
} orc_element;

typedef struct {
    double velocity[3];
    double force[3];
    double inertialMass;
} orc_node;

struct orc_context {
    snh_config cfg;
    snh_material mat[SNH_MAX_MATERIALS];
    unsigned nmat;
    unsigned I, J, K, EI, EJ, EK;
    unsigned N, E;
    double (*coord)[3];
    orc_node* node;
    orc_element* elem;
    unsigned (*elementNodeTbl)[8];
    unsigned (*nodeElementTbl)[8];
    double* temperature;           /* SnacTemperature_Node.temperature (zero
                                      unless the temperature plugin is on) */
    double* temperature0;          /* SnacTemperature_Node.temperature0 */
    double (*heatFlux)[NTET][3];   /* SnacTemperature_Element.heatFlux */
    double (*thermalStress)[NTET]; /* SnacTemperature_Element.thermalStress */
    int thermal, thermal_stress;
    unsigned char* tbc_mask;       /* temperature BC per node */
    double* tbc_value;
    double* refvisc_field;         /* per-element refvisc, or NULL */
    double* coh_offset_field;      /* per-element cohesion offset, or NULL */
    /* velocity BCs: per-node table and the index table of the composite */
    unsigned char* bc_mask;
    double (*bc_value)[3];
    unsigned* bc_index;            /* ascending node indices with a condition */
    unsigned bc_count;
    int bc_scan;
    /* context scalars */
    double dt, speedOfSound, minLengthScale, gravity, demf, strain_inert;
    double vmax;
    double topoGradMax;            /* running maximum of |grad y| on the top
                                      surface (Context.c:759-764), never reset */
    unsigned timeStep;
    int zero_done;
    /* optional record of the discrete decisions of each step (test
       infrastructure for the adjoint): sign(v) sign(f) of every node
       component before damping [N][3], plastic branch of every tet [E][10]
       (0 elastic, 1 shear, 2 tension) */
    int8_t* rec_damp;
    int8_t* rec_branch;
};

static __thread char g_err[256];
static int g_threads = 1;
static int g_glibc_hypot = 1;

/* hypot of tql2 (plastic_BI/Constitutive.c:430,462): the reference calls the
   system libm's, here glibc 2.35's, which is faithful but not correctly
   rounded (a few in 10^3 results are one ulp off).  The oracle calls the same
   function.  orc_set_glibc_hypot(0) switches to the correctly rounded value,
   obtained independently in binary128 (the 106-bit squares are exact, the
   sum and the square root are rounded to 113 bits, then once more to 53), to
   measure how much the choice matters. */
static double orc_hypot(double x, double y) {
    if (g_glibc_hypot) return hypot(x, y);
    if (isinf(x) || isinf(y)) return INFINITY;
    _Float128 s = (_Float128)x * x + (_Float128)y * y;
    return (double)sqrtf128(s);
}
void orc_set_glibc_hypot(int on) { g_glibc_hypot = on; }

static int fail(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return SNH_EINVAL;
}
const char* orc_last_error(void) { return g_err; }

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * x86-64 semantics of the reference's abs() on a double
 * (plastic_BI/Constitutive.c:290,409,413,483; viscoplastic_BI/Constitutive.c:
 * 441,560,564,634): the double is truncated to int by cvttsd2si, which yields
 * INT_MIN for NaN and out-of-range values, and abs(INT_MIN) stays INT_MIN.
 * ------------------------------------------------------------------------- */
int orc_ref_iabs(double x) {
    int t = (x > -2147483649.0 && x < 2147483648.0) ? (int)x : INT_MIN;
    if (t == INT_MIN) return INT_MIN;
    return t < 0 ? -t : t;
}
/* abs(a) + abs(b) in int, which wraps on x86 (tql2 tst1 update) */
static int iadd_wrap(int a, int b) { return (int)((unsigned)a + (unsigned)b); }

/* ---------------------------------------------------------------------------
 * Mesh construction
 * ------------------------------------------------------------------------- */
static void build_mesh(orc_context* c) {
    const unsigned I = c->I, J = c->J, K = c->K;
    const unsigned EI = c->EI, EJ = c->EJ;
    const double mn[3] = {c->cfg.minX, c->cfg.minY, c->cfg.minZ};
    const double mx[3] = {c->cfg.maxX, c->cfg.maxY, c->cfg.maxZ};
    const unsigned sz[3] = {I, J, K};
    /* BlockGeometry.c:326-349 */
    for (unsigned n = 0; n < c->N; n++) {
        unsigned ijk[3] = {n % I, (n / I) % J, n / (I * J)};
        for (int d = 0; d < 3; d++)
            c->coord[n][d] = sz[d] > 1 ? mn[d] + (mx[d] - mn[d]) * ((double)ijk[d] / (double)(sz[d] - 1)) : mn[d];
    }
    /* _HexaEL_BuildCornerIndices, HexaEL.c:456-470 */
#define P3(i, j, k) ((k) * I * J + (j) * I + (i))
    for (unsigned e = 0; e < c->E; e++) {
        unsigned i = e % EI, j = (e / EI) % EJ, k = e / (EI * EJ);
        c->elementNodeTbl[e][0] = P3(i, j, k);
        c->elementNodeTbl[e][1] = P3(i + 1, j, k);
        c->elementNodeTbl[e][3] = P3(i, j + 1, k);
        c->elementNodeTbl[e][2] = P3(i + 1, j + 1, k);
        c->elementNodeTbl[e][4] = P3(i, j, k + 1);
        c->elementNodeTbl[e][5] = P3(i + 1, j, k + 1);
        c->elementNodeTbl[e][7] = P3(i, j + 1, k + 1);
        c->elementNodeTbl[e][6] = P3(i + 1, j + 1, k + 1);
    }
#undef P3
    /* _HexaEL_BuildCornerElements3D, HexaEL.c:529-571: slot s = dx + 2 dy + 4 dz
       holds element (i-1+dx, j-1+dy, k-1+dz), sentinel = element count */
    for (unsigned n = 0; n < c->N; n++) {
        int i = (int)(n % I), j = (int)((n / I) % J), k = (int)(n / (I * J));
        for (unsigned s = 0; s < 8; s++) {
            int ei = i - 1 + (int)(s & 1), ej = j - 1 + (int)((s >> 1) & 1), ek = k - 1 + (int)((s >> 2) & 1);
            if (ei >= 0 && ej >= 0 && ek >= 0 && ei < (int)EI && ej < (int)EJ && ek < (int)c->EK)
                c->nodeElementTbl[n][s] = (unsigned)ek * EI * EJ + (unsigned)ej * EI + (unsigned)ei;
            else
                c->nodeElementTbl[n][s] = c->E;
        }
    }
}

static int node_on_wall(const orc_context* c, unsigned n, int wall) {
    unsigned i = n % c->I, j = (n / c->I) % c->J, k = n / (c->I * c->J);
    switch (wall) {
    case SNH_WALL_FRONT: return c->K > 1 && k == 0;            /* RegularMeshUtils.c */
    case SNH_WALL_BACK: return c->K > 1 && k == c->K - 1;
    case SNH_WALL_LEFT: return c->I > 1 && i == 0;
    case SNH_WALL_RIGHT: return c->I > 1 && i == c->I - 1;
    case SNH_WALL_TOP: return c->J > 1 && j == c->J - 1;
    case SNH_WALL_BOTTOM: return c->J > 1 && j == 0;
    case SNH_WALL_ALL: return 1;
    default: return 0;
    }
}

/* CompositeVC of WallVCs (CompositeVC.c:464-620): the index set is the union
   of the walls; per component the last item that constrains it wins. */
static void build_bcs(orc_context* c) {
    c->bc_count = 0;
    for (unsigned n = 0; n < c->N; n++) {
        unsigned char m = 0;
        int member = 0;
        for (unsigned b = 0; b < c->cfg.nbcs; b++) {
            const snh_velocity_bc* bc = &c->cfg.bcs[b];
            if (!node_on_wall(c, n, bc->wall)) continue;
            member = 1;
            for (int d = 0; d < 3; d++)
                if (bc->mask & (1u << d)) {
                    m |= (unsigned char)(1u << d);
                    c->bc_value[n][d] = bc->value[d];
                }
        }
        c->bc_mask[n] = m;
        if (member) c->bc_index[c->bc_count++] = n;
    }
}

/* VariableCondition_ApplyToIndex (VariableCondition.c:449-563) */
static void apply_bc(orc_context* c, unsigned n) {
    if (c->bc_scan) {
        unsigned idx;
        for (idx = 0; idx < c->bc_count; idx++)
            if (c->bc_index[idx] == n) break;
        if (idx == c->bc_count) return;
    }
    unsigned char m = c->bc_mask[n];
    for (int d = 0; d < 3; d++)
        if (m & (1u << d)) c->node[n].velocity[d] = c->bc_value[n][d];
}

orc_context* orc_create(const snh_config* cfg) {
    if (!cfg || cfg->meshSizeI < 2 || cfg->meshSizeJ < 2 || cfg->meshSizeK < 2) {
        fail("orc_create: need at least 2 nodes per axis");
        return NULL;
    }
    if (cfg->nmaterials < 1 || cfg->nmaterials > SNH_MAX_MATERIALS || !cfg->materials) {
        fail("orc_create: bad material list");
        return NULL;
    }
    orc_context* c = (orc_context*)calloc(1, sizeof *c);
    if (!c) return NULL;
    c->cfg = *cfg;
    c->nmat = cfg->nmaterials;
    memcpy(c->mat, cfg->materials, sizeof(snh_material) * cfg->nmaterials);
    c->cfg.bcs = NULL;   /* owned copy below; never the caller's array */
    if (cfg->nbcs) {
        snh_velocity_bc* b = (snh_velocity_bc*)malloc(sizeof(snh_velocity_bc) * cfg->nbcs);
        memcpy(b, cfg->bcs, sizeof(snh_velocity_bc) * cfg->nbcs);
        c->cfg.bcs = b;
    }
    c->cfg.materials = c->mat;
    c->I = cfg->meshSizeI; c->J = cfg->meshSizeJ; c->K = cfg->meshSizeK;
    c->EI = c->I - 1; c->EJ = c->J - 1; c->EK = c->K - 1;
    c->N = c->I * c->J * c->K;
    c->E = c->EI * c->EJ * c->EK;
    c->coord = calloc(c->N, sizeof *c->coord);
    c->node = calloc(c->N, sizeof *c->node);
    c->elem = calloc(c->E, sizeof *c->elem);   /* ExtensionManager.c:811-812 zeroes */
    c->elementNodeTbl = calloc(c->E, sizeof *c->elementNodeTbl);
    c->nodeElementTbl = calloc(c->N, sizeof *c->nodeElementTbl);
    c->temperature = calloc(c->N, sizeof(double));
    c->bc_mask = calloc(c->N, 1);
    c->bc_value = calloc(c->N, sizeof *c->bc_value);
    c->bc_index = calloc(c->N, sizeof(unsigned));
    if (!c->coord || !c->node || !c->elem || !c->elementNodeTbl || !c->nodeElementTbl || !c->temperature ||
        !c->bc_mask || !c->bc_value || !c->bc_index) {
        orc_destroy(c);
        fail("orc_create: out of memory");
        return NULL;
    }
    build_mesh(c);
    build_bcs(c);
    c->dt = cfg->timeStep;
    c->gravity = cfg->gravity;
    c->demf = cfg->demf;
    c->strain_inert = cfg->strain_inert;
    c->speedOfSound = 0.0;
    c->minLengthScale = 0.0;
    c->timeStep = 0;
    /* the initial velocity field is zero (nodeICs) with the BCs applied
       (_Snac_Context_InitialConditions, Context.c:772-773) */
    for (unsigned n = 0; n < c->N; n++) apply_bc(c, n);
    return c;
}

void orc_destroy(orc_context* c) {
    if (!c) return;
    free((void*)c->cfg.bcs);
    free(c->coord); free(c->node); free(c->elem);
    free(c->elementNodeTbl); free(c->nodeElementTbl);
    free(c->temperature); free(c->refvisc_field); free(c->coh_offset_field);
    free(c->temperature0); free(c->heatFlux); free(c->thermalStress); free(c->tbc_mask); free(c->tbc_value);
    free(c->bc_mask); free(c->bc_value); free(c->bc_index);
    free(c);
}

void orc_set_bc_scan(orc_context* c, int faithful) { c->bc_scan = faithful; }

void orc_record_decisions(orc_context* c, int8_t* damp, int8_t* branch) {
    c->rec_damp = damp;
    c->rec_branch = branch;
}

int orc_get_bc(orc_context* c, uint8_t* mask, double* value) {
    for (unsigned n = 0; n < c->N; n++) {
        mask[n] = c->bc_mask[n];
        for (int d = 0; d < 3; d++) value[3 * (size_t)n + d] = c->bc_value[n][d];
    }
    return SNH_OK;
}
uint32_t orc_node_count(orc_context* c) { return c->N; }
uint32_t orc_element_count(orc_context* c) { return c->E; }
int orc_set_dt(orc_context* c, double dt) { c->dt = dt; return SNH_OK; }

int orc_set_element_material(orc_context* c, const uint8_t* mat, size_t nelem) {
    if (nelem != c->E) return fail("element count mismatch");
    for (size_t e = 0; e < nelem; e++) {
        if (mat[e] >= c->nmat) return fail("material index out of range");
        c->elem[e].material_I = mat[e];
    }
    return SNH_OK;
}

int orc_set_element_field(orc_context* c, int field, const double* v, size_t nelem) {
    if (nelem != c->E) return fail("element count mismatch");
    double** dst = field == SNH_FIELD_REFVISC ? &c->refvisc_field
                 : field == SNH_FIELD_COHESION_OFFSET ? &c->coh_offset_field : NULL;
    if (!dst) return fail("unknown field");
    if (!*dst) *dst = malloc(sizeof(double) * c->E);
    memcpy(*dst, v, sizeof(double) * c->E);
    return SNH_OK;
}

int orc_set_node_temperature(orc_context* c, const double* t, size_t nnode) {
    if (nnode != c->N) return fail("node count mismatch");
    memcpy(c->temperature, t, sizeof(double) * c->N);
    return SNH_OK;
}

int orc_set_state(orc_context* c, const snh_state* s) {
    if (s->coord) memcpy(c->coord, s->coord, sizeof(double) * 3 * c->N);
    if (s->temperature) memcpy(c->temperature, s->temperature, sizeof(double) * c->N);
    if (s->velocity)
        for (unsigned n = 0; n < c->N; n++)
            for (int d = 0; d < 3; d++) c->node[n].velocity[d] = s->velocity[3 * n + d];
    for (unsigned e = 0; e < c->E; e++) {
        orc_element* el = &c->elem[e];
        for (int t = 0; t < NTET; t++) {
            if (s->stress) {
                const double* p = s->stress + ((size_t)e * NTET + t) * 6;
                double(*st)[3] = el->tetra[t].stress;
                memset(st, 0, sizeof el->tetra[t].stress);
                st[0][0] = p[0]; st[1][1] = p[1]; st[2][2] = p[2];
                st[0][1] = p[3]; st[0][2] = p[4]; st[1][2] = p[5];
            }
            if (s->plastic_strain) el->plasticStrain[t] = s->plastic_strain[(size_t)e * NTET + t];
        }
    }
    return SNH_OK;
}

int orc_get_state(orc_context* c, snh_state* s) {
    if (s->coord) memcpy(s->coord, c->coord, sizeof(double) * 3 * c->N);
    if (s->temperature) memcpy(s->temperature, c->temperature, sizeof(double) * c->N);
    if (s->velocity)
        for (unsigned n = 0; n < c->N; n++)
            for (int d = 0; d < 3; d++) s->velocity[3 * n + d] = c->node[n].velocity[d];
    for (unsigned e = 0; e < c->E; e++) {
        const orc_element* el = &c->elem[e];
        for (int t = 0; t < NTET; t++) {
            size_t k = (size_t)e * NTET + t;
            if (s->stress) {
                double* p = s->stress + k * 6;
                const double(*st)[3] = el->tetra[t].stress;
                p[0] = st[0][0]; p[1] = st[1][1]; p[2] = st[2][2];
                p[3] = st[0][1]; p[4] = st[0][2]; p[5] = st[1][2];
            }
            if (s->plastic_strain) s->plastic_strain[k] = el->plasticStrain[t];
            if (s->viscosity) s->viscosity[k] = el->viscosity[t];
            if (s->volume) s->volume[k] = el->tetra[t].volume;
        }
    }
    return SNH_OK;
}

int orc_get_scalars(orc_context* c, snh_scalars* out) {
    out->dt = c->dt;
    out->speedOfSound = c->speedOfSound;
    out->minLengthScale = c->minLengthScale;
    out->vmax = c->vmax;
    out->timeStep = c->timeStep;
    out->topoGradMax = c->topoGradMax;
    return SNH_OK;
}

int orc_element_nodes(orc_context* c, uint32_t* out) {
    memcpy(out, c->elementNodeTbl, sizeof(unsigned) * 8 * c->E);
    return SNH_OK;
}
int orc_node_elements(orc_context* c, uint32_t* out) {
    memcpy(out, c->nodeElementTbl, sizeof(unsigned) * 8 * c->N);
    return SNH_OK;
}

int orc_get_tet(orc_context* c, uint32_t e, uint32_t t, orc_tet_view* o) {
    if (e >= c->E || t >= NTET) return fail("index out of range");
    const orc_tet* tt = &c->elem[e].tetra[t];
    o->old_volume = tt->old_volume;
    o->volume = tt->volume;
    for (int s = 0; s < 4; s++) {
        o->area[s] = tt->surface[s].area;
        for (int d = 0; d < 3; d++) o->normal[s][d] = tt->surface[s].normal[d];
    }
    memcpy(o->strainRate, tt->strainRate, sizeof o->strainRate);
    memcpy(o->strain, tt->strain, sizeof o->strain);
    memcpy(o->stress, tt->stress, sizeof o->stress);
    o->density = tt->density;
    return SNH_OK;
}

int orc_get_element(orc_context* c, uint32_t e, orc_element_view* o) {
    if (e >= c->E) return fail("index out of range");
    o->volume = c->elem[e].volume;
    o->strainRate = c->elem[e].strainRate;
    o->stress = c->elem[e].stress;
    o->hydroPressure = c->elem[e].hydroPressure;
    o->aps = c->elem[e].aps;
    return SNH_OK;
}

int orc_get_node_force(orc_context* c, uint32_t n, double force[3], double* inertialMass) {
    if (n >= c->N) return fail("index out of range");
    for (int d = 0; d < 3; d++) force[d] = c->node[n].force[d];
    *inertialMass = c->node[n].inertialMass;
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * Tetrahedra.c:44-113 (re-entrant: locals instead of function statics)
 * ------------------------------------------------------------------------- */
static double tet_volume(const double* c1, const double* c2, const double* c3, const double* c4) {
    double x2 = c2[0] - c1[0], y2 = c2[1] - c1[1], z2 = c2[2] - c1[2];
    double x3 = c3[0] - c1[0], y3 = c3[1] - c1[1], z3 = c3[2] - c1[2];
    double x4 = c4[0] - c1[0], y4 = c4[1] - c1[1], z4 = c4[2] - c1[2];
    double v = (x4 * (y2 * z3 - z2 * y3) + y4 * (z2 * x3 - x2 * z3) + z4 * (x2 * y3 - x3 * y2)) / 6.0;
    return fabs(v);
}

static double tri_area(const double* c1, const double* c2, const double* c3) {
    double x2 = c2[0] - c1[0], y2 = c2[1] - c1[1], z2 = c2[2] - c1[2];
    double x3 = c3[0] - c1[0], y3 = c3[1] - c1[1], z3 = c3[2] - c1[2];
    double a = y2 * z3 - z2 * y3, b = z2 * x3 - x2 * z3, cc = x2 * y3 - y2 * x3;
    return sqrt(a * a + b * b + cc * cc) / 2.0;
}

static void tri_normal(const double* c1, const double* c2, const double* c3, double* n) {
    double x2 = c2[0] - c1[0], y2 = c2[1] - c1[1], z2 = c2[2] - c1[2];
    double x3 = c3[0] - c1[0], y3 = c3[1] - c1[1], z3 = c3[2] - c1[2];
    double ax = y2 * z3 - z2 * y3, ay = z2 * x3 - x2 * z3, az = x2 * y3 - y2 * x3;
    double aa = 1. * sqrt(ax * ax + ay * ay + az * az);
    n[0] = +1.0 * ax / aa;
    n[1] = +1.0 * ay / aa;
    n[2] = +1.0 * az / aa;
}

static const double* ecoord(const orc_context* c, unsigned e, unsigned k) {
    return c->coord[c->elementNodeTbl[e][k]];
}
static const double* evel(const orc_context* c, unsigned e, unsigned k) {
    return c->node[c->elementNodeTbl[e][k]].velocity;
}

/* ---------------------------------------------------------------------------
 * Snac_UpdateElementMomentum, UpdateElement.c:54-269
 * ------------------------------------------------------------------------- */
int orc_update_element(orc_context* c, uint32_t e, double* minLen) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    double rot[NTET][3];  /* [01], [02], [12] */
    if (c->cfg.dtType == SNH_DT_WAVE) {
        double Vp = sqrt((m->lambda + 2.0 * m->mu) / m->phsDensity);
        if (Vp > c->speedOfSound) c->speedOfSound = Vp;
    }
    el->volume = 0.0;
    *minLen = F_1EP27;
    for (int t = 0; t < NTET; t++) {
        orc_tet* tt = &el->tetra[t];
        const unsigned* tn = TetraToNode[t];
        memset(tt->strain, 0, sizeof tt->strain);
        tt->old_volume = tt->volume;
        tt->volume = tet_volume(ecoord(c, e, tn[0]), ecoord(c, e, tn[1]), ecoord(c, e, tn[2]), ecoord(c, e, tn[3]));
        if (c->timeStep == 0) tt->old_volume = tt->volume;
        for (int s = 0; s < 4; s++) {
            const double* p0 = ecoord(c, e, tn[FaceToNode[s][0]]);
            const double* p1 = ecoord(c, e, tn[FaceToNode[s][1]]);
            const double* p2 = ecoord(c, e, tn[FaceToNode[s][2]]);
            tt->surface[s].area = tri_area(p0, p1, p2);
            tri_normal(p0, p1, p2, tt->surface[s].normal);
        }
        for (int s = 0; s < 4; s++) {
            double ls = tt->surface[s].area;
            if (ls) {
                ls = fabs(tt->volume / ls);
                if (*minLen > ls) *minLen = ls;
            }
        }
        el->volume += tt->volume;

        /* rotation (:145-219): node k pairs with the face opposite it, (k+1)%4 */
        const double* v[4] = {evel(c, e, tn[0]), evel(c, e, tn[1]), evel(c, e, tn[2]), evel(c, e, tn[3])};
        const orc_surface* S = tt->surface;
#define TERM(k, a, b) (v[k][a] * S[((k) + 1) % 4].normal[b] * S[((k) + 1) % 4].area)
        double w01 = -(0.5 / tt->volume) *
                     (TERM(0, 0, 1) + TERM(1, 0, 1) + TERM(2, 0, 1) + TERM(3, 0, 1) - TERM(0, 1, 0) - TERM(1, 1, 0) -
                      TERM(2, 1, 0) - TERM(3, 1, 0)) / 3.0;
        double w02 = -(0.5 / tt->volume) *
                     (TERM(0, 0, 2) + TERM(1, 0, 2) + TERM(2, 0, 2) + TERM(3, 0, 2) - TERM(0, 2, 0) - TERM(1, 2, 0) -
                      TERM(2, 2, 0) - TERM(3, 2, 0)) / 3.0;
        double w12 = -(0.5 / tt->volume) *
                     (TERM(0, 1, 2) + TERM(1, 1, 2) + TERM(2, 1, 2) + TERM(3, 1, 2) - TERM(0, 2, 1) - TERM(1, 2, 1) -
                      TERM(2, 2, 1) - TERM(3, 2, 1)) / 3.0;
#undef TERM
        rot[t][0] = w01; rot[t][1] = w02; rot[t][2] = w12;
        /* strain rotation (:221-242) acts on the strain zeroed above: it stays zero */
        /* stress rotation, in place and in order (:244-265) */
        double(*st)[3] = tt->stress;
        const double dt = c->dt;
        st[0][0] += (2.0 * w01 * st[0][1] + 2.0 * w02 * st[0][2]) * dt;
        st[1][1] += (2.0 * (-1.0 * w01) * st[0][1] + 2.0 * w12 * st[1][2]) * dt;
        st[2][2] += (2.0 * (-1.0 * w02) * st[0][2] + 2.0 * (-1.0 * w12) * st[1][2]) * dt;
        st[0][1] += (w01 * (st[1][1] - st[0][0])) * dt + (w02 * st[1][2] + w12 * st[0][2]) * dt;
        st[0][2] += (w02 * (st[2][2] - st[0][0])) * dt + (w01 * st[1][2] - w12 * st[0][1]) * dt;
        st[1][2] += (w12 * (st[2][2] - st[1][1])) * dt + ((-1.0 * w01) * st[0][2] - w02 * st[0][1]) * dt;
    }
    (void)rot;
    el->volume *= 0.5;
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * Snac_StrainRate, StrainRate.c:49-305
 * ------------------------------------------------------------------------- */
int orc_strain_rate(orc_context* c, uint32_t e) {
    orc_element* el = &c->elem[e];
    double trace[NTET];
    double partial, srVolAvg, srOtherAvg;
    const double elemVolume = 2.0 * el->volume;
    for (int t = 0; t < NTET; t++) {
        orc_tet* tt = &el->tetra[t];
        const unsigned* tn = TetraToNode[t];
        const double* v[4] = {evel(c, e, tn[0]), evel(c, e, tn[1]), evel(c, e, tn[2]), evel(c, e, tn[3])};
        const orc_surface* S = tt->surface;
        double(*sr)[3] = tt->strainRate;
        memset(sr, 0, sizeof tt->strainRate);
#define TERM(k, a, b) (v[k][a] * S[((k) + 1) % 4].normal[b] * S[((k) + 1) % 4].area)
#define SUM4(a, b) (TERM(0, a, b) + TERM(1, a, b) + TERM(2, a, b) + TERM(3, a, b))
        sr[0][0] = -1.0 * SUM4(0, 0) / 3.0 / tt->volume;
        sr[1][1] = -1.0 * SUM4(1, 1) / 3.0 / tt->volume;
        sr[2][2] = -1.0 * SUM4(2, 2) / 3.0 / tt->volume;
        sr[0][1] = -0.5 * (TERM(0, 0, 1) + TERM(1, 0, 1) + TERM(2, 0, 1) + TERM(3, 0, 1) + TERM(0, 1, 0) +
                           TERM(1, 1, 0) + TERM(2, 1, 0) + TERM(3, 1, 0)) / 3.0 / tt->volume;
        sr[0][2] = -0.5 * (TERM(0, 0, 2) + TERM(1, 0, 2) + TERM(2, 0, 2) + TERM(3, 0, 2) + TERM(0, 2, 0) +
                           TERM(1, 2, 0) + TERM(2, 2, 0) + TERM(3, 2, 0)) / 3.0 / tt->volume;
        sr[1][2] = -0.5 * (TERM(0, 1, 2) + TERM(1, 1, 2) + TERM(2, 1, 2) + TERM(3, 1, 2) + TERM(0, 2, 1) +
                           TERM(1, 2, 1) + TERM(2, 2, 1) + TERM(3, 2, 1)) / 3.0 / tt->volume;
#undef SUM4
#undef TERM
    }
    /* mixed discretisation (:272-289) */
    partial = 0.0;
    for (int t = 0; t < NTET; t++) {
        const double(*sr)[3] = el->tetra[t].strainRate;
        trace[t] = (sr[0][0] + sr[1][1] + sr[2][2]) / 3.0;
        partial += trace[t] * el->tetra[t].volume;
    }
    srVolAvg = 0.0;
    srOtherAvg = 0.0;
    for (int t = 0; t < NTET; t++) {
        double(*sr)[3] = el->tetra[t].strainRate;
        sr[0][0] += -1.0 * trace[t] + partial / elemVolume;
        sr[1][1] += -1.0 * trace[t] + partial / elemVolume;
        sr[2][2] += -1.0 * trace[t] + partial / elemVolume;
        srVolAvg += sr[1][1] * sr[2][2] + sr[2][2] * sr[0][0] + sr[0][0] * sr[1][1];
        srOtherAvg += sr[0][1] * sr[0][1] + sr[1][2] * sr[1][2] + sr[0][2] * sr[0][2];
    }
    srVolAvg /= NTET;
    srOtherAvg /= NTET;
    el->strainRate = 0.5 * sqrt(0.5 * fabs(-1.0 * srVolAvg + srOtherAvg));
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * plastic_BI eigen-solver: principal_stresses / tred2 / tql2
 * (Snac/plugins/plastic_BI/Constitutive.c:240-510; the EISPACK/JAMA
 * Householder tridiagonalisation and implicit QL, n = 3), with the
 * reference's int abs() and int-sum semantics.
 * ------------------------------------------------------------------------- */
static void tred2_3(double d[3], double e[3], double V[3][3]) {
    const int n = 3;
    for (int j = 0; j < n; j++) d[j] = V[n - 1][j];
    for (int i = n - 1; i > 0; i--) {
        double scale = 0.0, h = 0.0;
        for (int k = 0; k < i; k++) scale = scale + orc_ref_iabs(d[k]);
        if (scale == 0.0) {
            e[i] = d[i - 1];
            for (int j = 0; j < i; j++) {
                d[j] = V[i - 1][j];
                V[i][j] = 0.0;
                V[j][i] = 0.0;
            }
        } else {
            double f, g, hh;
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
            for (int j = 0; j < i; j++) e[j] = 0.0;
            for (int j = 0; j < i; j++) {
                f = d[j];
                V[j][i] = f;
                g = e[j] + V[j][j] * f;
                for (int k = j + 1; k <= i - 1; k++) {
                    g += V[k][j] * d[k];
                    e[k] += V[k][j] * f;
                }
                e[j] = g;
            }
            f = 0.0;
            for (int j = 0; j < i; j++) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            hh = f / (h + h);
            for (int j = 0; j < i; j++) e[j] -= hh * d[j];
            for (int j = 0; j < i; j++) {
                f = d[j];
                g = e[j];
                for (int k = j; k <= i - 1; k++) V[k][j] -= (f * e[k] + g * d[k]);
                d[j] = V[i - 1][j];
                V[i][j] = 0.0;
            }
        }
        d[i] = h;
    }
    for (int i = 0; i < n - 1; i++) {
        double h, g;
        V[n - 1][i] = V[i][i];
        V[i][i] = 1.0;
        h = d[i + 1];
        if (h != 0.0) {
            for (int k = 0; k <= i; k++) d[k] = V[k][i + 1] / h;
            for (int j = 0; j <= i; j++) {
                g = 0.0;
                for (int k = 0; k <= i; k++) g += V[k][i + 1] * V[k][j];
                for (int k = 0; k <= i; k++) V[k][j] -= g * d[k];
            }
        }
        for (int k = 0; k <= i; k++) V[k][i + 1] = 0.0;
    }
    for (int j = 0; j < n; j++) {
        d[j] = V[n - 1][j];
        V[n - 1][j] = 0.0;
    }
    V[n - 1][n - 1] = 1.0;
    e[0] = 0.0;
}

static void tql2_3(double d[3], double e[3], double V[3][3]) {
    const int n = 3;
    double f = 0.0, tst1 = 0.0;
    const double eps = 2.220446049250313080847e-16; /* pow(2.0,-52.0) */
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int m = l;
        {   /* tst1 = MAX(tst1, abs(d[l]) + abs(e[l])) with the int sum */
            double b = (double)iadd_wrap(orc_ref_iabs(d[l]), orc_ref_iabs(e[l]));
            tst1 = (tst1 > b) ? tst1 : b;
        }
        while (m < n) {
            if (orc_ref_iabs(e[m]) <= eps * tst1) break;
            m++;
        }
        if (m > l) {
            do {
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = orc_hypot(p, 1.0);
                double dl1, h, cc, c2, c3, el1, s, s2;
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                dl1 = d[l + 1];
                h = g - d[l];
                for (int i = l + 2; i < n; i++) d[i] -= h;
                f = f + h;
                p = d[m];
                cc = 1.0; c2 = cc; c3 = cc;
                el1 = e[l + 1];
                s = 0.0; s2 = 0.0;
                for (int i = m - 1; i >= l; i--) {
                    c3 = c2;
                    c2 = cc;
                    s2 = s;
                    g = cc * e[i];
                    h = cc * p;
                    r = orc_hypot(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    cc = p / r;
                    p = cc * d[i] - s * g;
                    d[i + 1] = h + s * (cc * g + s * d[i]);
                    for (int k = 0; k < n; k++) {
                        h = V[k][i + 1];
                        V[k][i + 1] = s * V[k][i] + cc * h;
                        V[k][i] = cc * V[k][i] - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = cc * p;
            } while (orc_ref_iabs(e[l]) > eps * tst1);
        }
        d[l] = d[l] + f;
        e[l] = 0.0;
    }
    for (int i = 0; i < n - 1; i++) {
        double p = d[i];
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; j++) {
                p = V[j][i];
                V[j][i] = V[j][k];
                V[j][k] = p;
            }
        }
    }
}

void orc_principal_stresses(const double stress[3][3], double d[3], double V[3][3]) {
    double e[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) V[i][j] = stress[i][j];
    tred2_3(d, e, V);
    tql2_3(d, e, V);
}

/* piecewise-linear softening (plastic_BI/Constitutive.c:116-132;
   viscoplastic_BI/Constitutive.c:266-280, which also returns tgc) */
static void soften(const snh_material* m, double ps, double* phi, double* psi, double* coh, double* hard) {
    for (unsigned i = 0; i < m->nsegments; i++) {
        const double pl1 = m->plstrain[i], pl2 = m->plstrain[i + 1];
        if (ps >= pl1 && ps <= pl2) {
            const double tgf = (m->frictionAngle[i + 1] - m->frictionAngle[i]) / (pl2 - pl1);
            const double tgd = (m->dilationAngle[i + 1] - m->dilationAngle[i]) / (pl2 - pl1);
            const double tgc = (m->cohesion[i + 1] - m->cohesion[i]) / (pl2 - pl1);
            *phi = m->frictionAngle[i] + tgf * (ps - pl1);
            *psi = m->dilationAngle[i] + tgd * (ps - pl1);
            *coh = m->cohesion[i] + tgc * (ps - pl1);
            if (hard) *hard = tgc;
        }
    }
}

/* Mohr-Coulomb return in principal space and back-rotation; shared by
   plastic_BI (:159-223, hardening = 0) and viscoplastic_BI (:307-371). */
static void mc_return(const snh_material* m, double (*st)[3], double s[3], double cn[3][3], double frictionAngle,
                      double dilationAngle, double cohesion, double tension_cutoff, double hardening, int with_hard,
                      double* plasticStrain, int8_t* branch) {
    const double degrad = M_PI / 180.0;
    const double a1 = m->lambda + 2.0 * m->mu;
    const double a2 = m->lambda;
    double alam, dep1, dep2, dep3, depm;
    double sphi = sin(frictionAngle * degrad);
    double spsi = sin(dilationAngle * degrad);
    double anphi = (1.0 + sphi) / (1.0 - sphi);
    double anpsi = (1.0 + spsi) / (1.0 - spsi);
    double fs = s[0] - s[2] * anphi + 2 * cohesion * sqrt(anphi);
    double ft = s[2] - tension_cutoff;
    int ind = 0;
    if (fs < 0.0 || ft > 0.0) {
        double aP = sqrt(1.0 + anphi * anphi) + anphi;
        double sP = tension_cutoff * anphi - 2 * cohesion * sqrt(anphi);
        double h = s[2] - tension_cutoff + aP * (s[0] - sP);
        ind = 1;
        if (branch) *branch = h < 0.0 ? 1 : 2;
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
    } else {
        dep1 = 0.0;
        dep2 = 0.0;
        dep3 = 0.0;
        if (branch) *branch = 0;
    }
    if (ind) {
        depm = (dep1 + dep2 + dep3) / 3.0;
        *plasticStrain += sqrt(0.5 * ((dep1 - depm) * (dep1 - depm) + (dep2 - depm) * (dep2 - depm) +
                                      (dep3 - depm) * (dep3 - depm) + depm * depm));
        memset(st, 0, sizeof(double) * 9);
        for (int mm = 0; mm < 3; mm++)
            for (int nn = mm; nn < 3; nn++)
                for (int k = 0; k < 3; k++) st[mm][nn] += cn[mm][k] * cn[nn][k] * s[k];
    }
}

/* SnacPlastic_Constitutive, plastic_BI/Constitutive.c:59-238.  cohesion,
   frictionAngle, dilationAngle and tension_cutoff are declared once per element
   (:78-81), so a tet whose plastic strain lies outside every segment keeps the
   previous tet's values. */
static int constitutive_plastic(orc_context* c, unsigned e) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    const double degrad = M_PI / 180.0;
    double totalVolume = 0.0, depls = 0.0;
    double cohesion = 30.0e6, frictionAngle = 30.0, dilationAngle = 0.0, tension_cutoff = 0.0;
    const double coh_off = c->coh_offset_field ? c->coh_offset_field[e] : 0.0;
    for (int t = 0; t < NTET; t++) {
        double cn[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double s[3] = {0.0, 0.0, 0.0};
        double(*st)[3] = el->tetra[t].stress;
        const double(*sn)[3] = el->tetra[t].strain;
        double plasticStrain = el->plasticStrain[t];
        double trace_strain = sn[0][0] + sn[1][1] + sn[2][2];
        double hard = 0.0, coh;
        st[0][0] += (2.0 * m->mu) * sn[0][0] + m->lambda * (trace_strain);
        st[1][1] += (2.0 * m->mu) * sn[1][1] + m->lambda * (trace_strain);
        st[2][2] += (2.0 * m->mu) * sn[2][2] + m->lambda * (trace_strain);
        st[0][1] += (2.0 * m->mu) * sn[0][1];
        st[0][2] += (2.0 * m->mu) * sn[0][2];
        st[1][2] += (2.0 * m->mu) * sn[1][2];
        st[1][0] = st[0][1];
        st[2][0] = st[0][2];
        st[2][1] = st[1][2];
        orc_principal_stresses((const double(*)[3])st, s, cn);
        soften(m, plasticStrain, &frictionAngle, &dilationAngle, &cohesion, &hard);
        coh = c->coh_offset_field ? cohesion + coh_off : cohesion;
        if (frictionAngle >= 0.0) {
            tension_cutoff = m->ten_off;
            if (frictionAngle > 0.0 && tension_cutoff == 0.0) tension_cutoff = coh / tan(frictionAngle * degrad);
        } else if (plasticStrain < 0.0) {
            el->plasticStrain[t] = 0.0;
            plasticStrain = 0.0;
            frictionAngle = m->frictionAngle[0];
            dilationAngle = m->dilationAngle[0];
            cohesion = m->cohesion[0];
            coh = c->coh_offset_field ? cohesion + coh_off : cohesion;
        } else {
            snprintf(g_err, sizeof g_err, "plastic: negative friction angle (elem %u tet %d)", e, t);
            return SNH_ENUMERIC;
        }
        mc_return(m, st, s, cn, frictionAngle, dilationAngle, coh, tension_cutoff, 0.0, 0, &el->plasticStrain[t],
                  c->rec_branch ? &c->rec_branch[(size_t)e * NTET + t] : NULL);
        depls += el->plasticStrain[t] * el->tetra[t].volume;
        totalVolume += el->tetra[t].volume;
    }
    el->aps = depls / totalVolume;
    return SNH_OK;
}

/* _SnacElastic_Constitutive, elastic/Constitutive.c:40-66 */
static void constitutive_elastic(orc_context* c, unsigned e) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    for (int t = 0; t < NTET; t++) {
        const double a2 = m->lambda;
        const double sh2 = 2.0 * m->mu;
        double(*st)[3] = el->tetra[t].stress;
        const double(*sn)[3] = el->tetra[t].strain;
        double trace_strain = sn[0][0] + sn[1][1] + sn[2][2];
        st[0][0] += sh2 * sn[0][0] + a2 * (trace_strain);
        st[1][1] += sh2 * sn[1][1] + a2 * (trace_strain);
        st[2][2] += sh2 * sn[2][2] + a2 * (trace_strain);
        st[0][1] += sh2 * sn[0][1];
        st[0][2] += sh2 * sn[0][2];
        st[1][2] += sh2 * sn[1][2];
    }
}

/* viscosity law shared by maxwell (:123-151) and viscoplastic_BI (:174-218);
   the temperature branch is taken because Snac_EP_LoopElementsEnergy always
   exists (Context.c:498-501) */
static double viscosity_of(orc_context* c, unsigned e, int t, const snh_material* m, double sd0, double sd1,
                           double sd2, const double (*sn)[3]) {
    const double R = 8.31448;
    double rviscosity = c->refvisc_field ? c->refvisc_field[e] : m->refvisc;
    double rstrainrate = m->refsrate, rTemp = m->reftemp, H = m->activationE, srexponent = m->srexponent;
    double srJ2 = sqrt(fabs(sd1 * sd2 + sd2 * sd0 + sd0 * sd1 - sn[0][1] * sn[0][1] - sn[0][2] * sn[0][2] -
                            sn[1][2] * sn[1][2])) /
                  c->dt;
    if (srJ2 == 0.0) srJ2 = rstrainrate;
    double avgTemp = 0.0;
    for (int k = 0; k < 4; k++) avgTemp += 0.25 * c->temperature[c->elementNodeTbl[e][TetraToNode[t][k]]];
    double visc = rviscosity * pow((srJ2 / rstrainrate), (1. / srexponent - 1.)) *
                  exp(H / R * (1. / (avgTemp + 273.15) - 1. / (rTemp + 273.15)));
    if (visc < m->vis_min) visc = m->vis_min;
    if (visc > m->vis_max) visc = m->vis_max;
    return visc;
}

/* _SnacMaxwell_Constitutive, maxwell/Constitutive.c:47-184 */
static int constitutive_maxwell(orc_context* c, unsigned e) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    for (int t = 0; t < NTET; t++) {
        const double bulkm = m->lambda + 2.0 * m->mu / 3.0;
        double(*st)[3] = el->tetra[t].stress;
        const double(*sn)[3] = el->tetra[t].strain;
        double rmu = m->mu;
        if (c->thermal_stress) {   /* :106-110 */
            st[0][0] += c->thermalStress[e][t];
            st[1][1] += c->thermalStress[e][t];
            st[2][2] += c->thermalStress[e][t];
        }
        double trace_stress = st[0][0] + st[1][1] + st[2][2];
        double trace_strain = el->tetra[t].volume / el->tetra[t].old_volume - 1.0;
        double straind0 = sn[0][0] - (trace_strain) / 3.0;
        double straind1 = sn[1][1] - (trace_strain) / 3.0;
        double straind2 = sn[2][2] - (trace_strain) / 3.0;
        double stressd0 = st[0][0] - (trace_stress) / 3.0;
        double stressd1 = st[1][1] - (trace_stress) / 3.0;
        double stressd2 = st[2][2] - (trace_stress) / 3.0;
        double visc = viscosity_of(c, e, t, m, straind0, straind1, straind2, sn);
        if (isnan(visc) || isinf(visc)) {
            snprintf(g_err, sizeof g_err, "maxwell: viscosity not finite (elem %u tet %d)", e, t);
            return SNH_ENUMERIC;
        }
        el->viscosity[t] = visc;
        double temp = rmu / (2.0 * visc) * c->dt;
        double vic1 = 1.0 - temp;
        double vic2 = 1.0 / (1.0 + temp);
        stressd0 = (stressd0 * vic1 + 2.0 * rmu * straind0) * vic2;
        stressd1 = (stressd1 * vic1 + 2.0 * rmu * straind1) * vic2;
        stressd2 = (stressd2 * vic1 + 2.0 * rmu * straind2) * vic2;
        st[0][1] = (st[0][1] * vic1 + 2.0 * rmu * sn[0][1]) * vic2;
        st[0][2] = (st[0][2] * vic1 + 2.0 * rmu * sn[0][2]) * vic2;
        st[1][2] = (st[1][2] * vic1 + 2.0 * rmu * sn[1][2]) * vic2;
        double VolumicStress = trace_stress / 3.0 + bulkm * trace_strain;
        st[0][0] = stressd0 + VolumicStress;
        st[1][1] = stressd1 + VolumicStress;
        st[2][2] = stressd2 + VolumicStress;
    }
    return SNH_OK;
}

/* SnacViscoPlastic_Constitutive, viscoplastic_BI/Constitutive.c:60-389; the
   softened parameters persist across the tets of an element (:115-119). */
static int constitutive_viscoplastic(orc_context* c, unsigned e) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    const double degrad = M_PI / 180.0;
    const double bulkm = m->lambda + 2.0 * m->mu / 3.0;
    const double rmu = m->mu;
    double totalVolume = 0.0, depls = 0.0;
    double cohesion = 0.0, frictionAngle = 0.0, dilationAngle = 0.0, hardening = 0.0, tension_cutoff = 0.0;
    const double coh_off = c->coh_offset_field ? c->coh_offset_field[e] : 0.0;
    for (int t = 0; t < NTET; t++) {
        double cn[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double s[3] = {0.0, 0.0, 0.0};
        double(*st)[3] = el->tetra[t].stress;
        const double(*sn)[3] = el->tetra[t].strain;
        double plasticStrain = el->plasticStrain[t];
        if (c->thermal_stress) {   /* :145-149 */
            st[0][0] += c->thermalStress[e][t];
            st[1][1] += c->thermalStress[e][t];
            st[2][2] += c->thermalStress[e][t];
        }
        double trace_stress = st[0][0] + st[1][1] + st[2][2];
        double trace_strain = sn[0][0] + sn[1][1] + sn[2][2];
        double straind0 = sn[0][0] - (trace_strain) / 3.0;
        double straind1 = sn[1][1] - (trace_strain) / 3.0;
        double straind2 = sn[2][2] - (trace_strain) / 3.0;
        double stressd0 = st[0][0] - (trace_stress) / 3.0;
        double stressd1 = st[1][1] - (trace_stress) / 3.0;
        double stressd2 = st[2][2] - (trace_stress) / 3.0;
        double coh;
        double visc = viscosity_of(c, e, t, m, straind0, straind1, straind2, sn);
        if (isnan(visc) || isinf(visc)) {
            snprintf(g_err, sizeof g_err, "viscoplastic: viscosity not finite (elem %u tet %d)", e, t);
            return SNH_ENUMERIC;
        }
        el->viscosity[t] = visc;
        double temp = rmu / (2.0 * visc) * c->dt;
        double vic1 = 1.0 - temp;
        double vic2 = 1.0 / (1.0 + temp);
        stressd0 = (stressd0 * vic1 + 2.0 * rmu * straind0) * vic2;
        stressd1 = (stressd1 * vic1 + 2.0 * rmu * straind1) * vic2;
        stressd2 = (stressd2 * vic1 + 2.0 * rmu * straind2) * vic2;
        st[0][1] = (st[0][1] * vic1 + 2.0 * rmu * sn[0][1]) * vic2;
        st[0][2] = (st[0][2] * vic1 + 2.0 * rmu * sn[0][2]) * vic2;
        st[1][2] = (st[1][2] * vic1 + 2.0 * rmu * sn[1][2]) * vic2;
        double VolumicStress = trace_stress / 3.0 + bulkm * trace_strain;
        st[0][0] = stressd0 + VolumicStress;
        st[1][1] = stressd1 + VolumicStress;
        st[2][2] = stressd2 + VolumicStress;
        /* not symmetrised: the eigen-solve reads the lower triangle, which
           stays zero (:246) */
        orc_principal_stresses((const double(*)[3])st, s, cn);
        soften(m, plasticStrain, &frictionAngle, &dilationAngle, &cohesion, &hardening);
        coh = c->coh_offset_field ? cohesion + coh_off : cohesion;
        if (frictionAngle >= 0.0) {
            tension_cutoff = m->ten_off;
            if (frictionAngle > 0.0 && tension_cutoff == 0.0) tension_cutoff = coh / tan(frictionAngle * degrad);
        } else if (plasticStrain < 0.0) {
            el->plasticStrain[t] = 0.0;
            plasticStrain = 0.0;
            frictionAngle = m->frictionAngle[0];
            dilationAngle = m->dilationAngle[0];
            cohesion = m->cohesion[0];
            coh = c->coh_offset_field ? cohesion + coh_off : cohesion;
        } else {
            snprintf(g_err, sizeof g_err, "viscoplastic: negative friction angle (elem %u tet %d)", e, t);
            return SNH_ENUMERIC;
        }
        mc_return(m, st, s, cn, frictionAngle, dilationAngle, coh, tension_cutoff, hardening, 1,
                  &el->plasticStrain[t], NULL);
        depls += el->plasticStrain[t] * el->tetra[t].volume;
        totalVolume += el->tetra[t].volume;
    }
    el->aps = depls / totalVolume;
    return SNH_OK;
}

/* Snac_EP_Constitutive hooks in plugin order: each acts only on materials
   carrying its rheology bit (one rheology plugin per run in the configs). */
static int constitutive(orc_context* c, unsigned e) {
    const snh_material* m = &c->mat[c->elem[e].material_I];
    int rc = SNH_OK;
    if (m->rheology & SNH_RHEOLOGY_PLASTIC) {
        rc = constitutive_plastic(c, e);
    } else if (m->rheology & SNH_RHEOLOGY_VISCOPLASTIC) {
        rc = constitutive_viscoplastic(c, e);
    } else if (m->rheology & SNH_RHEOLOGY_MAXWELL) {
        rc = constitutive_maxwell(c, e);
    } else if (m->rheology & SNH_RHEOLOGY_ELASTIC) {
        constitutive_elastic(c, e);
    }
    return rc;
}

/* stress trace averaging and element scalars (Stress.c:87-171) */
static void stress_average(orc_context* c, unsigned e, double alpha, double beta) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    const double elemVolume = 2.0 * el->volume;
    double trace[NTET], partial = 0.0, sVolAvg = 0.0, sOtherAvg = 0.0, pressure = 0.0;
    for (int t = 0; t < NTET; t++) {
        const double(*st)[3] = el->tetra[t].stress;
        trace[t] = (st[0][0] + st[1][1] + st[2][2]) / 3.0;
        partial += trace[t] * el->tetra[t].volume;
    }
    for (int t = 0; t < NTET; t++) {
        const double averageTraceStress = partial / elemVolume;
        double(*st)[3] = el->tetra[t].stress;
        st[0][0] += -1.0 * trace[t] + averageTraceStress;
        st[1][1] += -1.0 * trace[t] + averageTraceStress;
        st[2][2] += -1.0 * trace[t] + averageTraceStress;
        sVolAvg += (st[1][1] - averageTraceStress) * (st[2][2] - averageTraceStress) +
                   (st[2][2] - averageTraceStress) * (st[0][0] - averageTraceStress) +
                   (st[0][0] - averageTraceStress) * (st[1][1] - averageTraceStress);
        sOtherAvg += st[0][1] * st[0][1] + st[1][2] * st[1][2] + st[0][2] * st[0][2];
        pressure += (st[0][0] + st[1][1] + st[2][2]) / 3.0;
    }
    sVolAvg /= NTET;
    sOtherAvg /= NTET;
    pressure /= NTET;
    el->stress = 0.5 * sqrt(0.5 * fabs(-1.0 * sVolAvg + sOtherAvg));
    el->hydroPressure = -1.0 * pressure;
    for (int t = 0; t < NTET; t++)
        el->tetra[t].density =
            m->phsDensity * (1.0 - alpha * (el->tetra[t].avgTemp - m->reftemp) + beta * pressure);
}

/* Snac_Stress, Stress.c:53-172 */
int orc_stress(orc_context* c, uint32_t e) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    for (int t = 0; t < NTET; t++) {
        orc_tet* tt = &el->tetra[t];
        memset(tt->strain, 0, sizeof tt->strain);
        tt->strain[0][0] = tt->strainRate[0][0] * c->dt;
        tt->strain[1][1] = tt->strainRate[1][1] * c->dt;
        tt->strain[2][2] = tt->strainRate[2][2] * c->dt;
        tt->strain[0][1] = tt->strainRate[0][1] * c->dt;
        tt->strain[0][2] = tt->strainRate[0][2] * c->dt;
        tt->strain[1][2] = tt->strainRate[1][2] * c->dt;
    }
    int rc = constitutive(c, e);
    if (rc) return rc;
    stress_average(c, e, m->alpha, m->beta);
    if (isnan(el->stress) || isinf(el->stress)) {   /* Journal_OFirewall, Stress.c:161-166 */
        snprintf(g_err, sizeof g_err, "timeStep=%u Element %u: stress scalar not finite", c->timeStep, e);
        return SNH_ENUMERIC;
    }
    return SNH_OK;
}

/* SnacTest_Stress (Snac/libSnac/tests/TestStress.c): inline elastic law with
   alpha = beta = 0; used by the reference's node-update unit tests. */
int orc_test_stress(orc_context* c, uint32_t e) {
    orc_element* el = &c->elem[e];
    for (int t = 0; t < NTET; t++) {
        orc_tet* tt = &el->tetra[t];
        memset(tt->strain, 0, sizeof tt->strain);
        tt->strain[0][0] = tt->strainRate[0][0] * c->dt;
        tt->strain[1][1] = tt->strainRate[1][1] * c->dt;
        tt->strain[2][2] = tt->strainRate[2][2] * c->dt;
        tt->strain[0][1] = tt->strainRate[0][1] * c->dt;
        tt->strain[0][2] = tt->strainRate[0][2] * c->dt;
        tt->strain[1][2] = tt->strainRate[1][2] * c->dt;
    }
    constitutive_elastic(c, e);
    stress_average(c, e, 0.0, 0.0);
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * Snac_Force, Force.c:53-188 ("complete" force calculation, Cartesian)
 * ------------------------------------------------------------------------- */
int orc_force(orc_context* c, uint32_t n, double speedOfSound, double* mass, double* inertialMass, double force[3],
              double balance[3]) {
    const double factor4 = 1.0 / 4.0;
    const double gravity = c->gravity;
    *inertialMass = 0.0;
    *mass = 0.0;
    force[0] = force[1] = force[2] = 0.0;
    balance[0] = balance[1] = balance[2] = 0.0;
    for (unsigned ne = 0; ne < 8; ne++) {
        unsigned e = c->nodeElementTbl[n][ne];
        if (e >= c->E) continue;
        const orc_element* el = &c->elem[e];
        const snh_material* m = &c->mat[el->material_I];
        const double inertialDensity = (m->lambda + 2.0 * m->mu) / (speedOfSound * speedOfSound);
        for (int idx = 0; idx < 5; idx++) {
            const unsigned t = NodeToTetra[ne][idx];
            const unsigned s = NodeToSurface[ne][idx];
            const orc_tet* tt = &el->tetra[t];
            const double(*st)[3] = tt->stress;
            const orc_surface* sf = &tt->surface[s];
            const double* nrm = sf->normal;
            const double effDensity = tt->density;
            if (c->cfg.dtType == SNH_DT_DYNAMIC) {
                *inertialMass += factor4 * inertialDensity * tt->volume;
            } else if (c->cfg.dtType == SNH_DT_CONSTANT) {
                double alpha1 = (m->lambda + 2.0 * m->mu), area_sum = 0.0, temp;
                for (int d = 0; d < 3; d++) {
                    temp = (nrm[d] * sf->area) * (nrm[d] * sf->area);
                    if (temp > area_sum) area_sum = temp;
                }
                *inertialMass += alpha1 * area_sum / (9.0 * tt->volume) * (c->dt * c->dt);
            } else {
                *inertialMass += factor4 * effDensity * tt->volume;
            }
            force[0] += (st[0][0] * nrm[0] + st[0][1] * nrm[1] + st[0][2] * nrm[2]) * sf->area / 6.0;
            force[1] += (st[0][1] * nrm[0] + st[1][1] * nrm[1] + st[1][2] * nrm[2]) * sf->area / 6.0;
            force[2] += (st[0][2] * nrm[0] + st[1][2] * nrm[1] + st[2][2] * nrm[2]) * sf->area / 6.0;
            force[1] -= gravity * effDensity * tt->volume / 8.0;
            balance[0] += fabs(force[0]);
            balance[1] += fabs(force[1]);
            balance[2] += fabs(force[2]);
            *mass += effDensity * tt->volume / 8.0;
        }
    }
    return SNH_OK;
}

#define GETSIGN(val) ((val) < 0 ? -1.0 : (val) > 0 ? +1.0 : 0.0)

/* Snac_UpdateNodeMomentum_PreProcess, UpdateNode.c:74-158 (Cartesian branch) */
int orc_update_node_preprocess(orc_context* c, uint32_t n, double inertialMass, double force[3]) {
    orc_node* nd = &c->node[n];
    if (c->rec_damp)
        for (int d = 0; d < 3; d++) c->rec_damp[3 * (size_t)n + d] = (int8_t)(GETSIGN(nd->velocity[d]) * GETSIGN(force[d]));
    force[0] -= c->demf * GETSIGN(nd->velocity[0]) * fabs(force[0]);
    force[1] -= c->demf * GETSIGN(nd->velocity[1]) * fabs(force[1]);
    force[2] -= c->demf * GETSIGN(nd->velocity[2]) * fabs(force[2]);
    nd->velocity[0] += c->dt * force[0] / inertialMass;
    nd->velocity[1] += c->dt * force[1] / inertialMass;
    nd->velocity[2] += c->dt * force[2] / inertialMass;
    apply_bc(c, n);
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * Topography diffusion of the top surface (topo_kappa > 0), Cartesian meshes.
 * More than two nodes wide in x and z the surface is diffused in two
 * dimensions; two nodes wide in x (resp. z) and more in the other direction
 * it is diffused along z (resp. x) only, as diffTopo_alongZ/_alongX,
 * getMaxTopoGrad_alongZ/_alongX and AdjustEdgeTopo's last two branches do
 * (UpdateNode.c:248-251,262-357,494-499,508-645,686-708,811-856); two nodes
 * wide in both it is left alone.  The reference runs it per node inside
 * the second node pass, in node order, after the node's coordinate update:
 * neighbours with a smaller index already hold their new coordinates (and
 * diffused heights), those with a larger one still hold the old ones.  The
 * neighbour slots of IJK6Topology_BuildNeighbours (StGermain/Discretisation/
 * Geometry/src/IJK6Topology.c:283-343, fixed positions): 0 +x, 2 +z, 3 -x,
 * 5 -z.
 * ------------------------------------------------------------------------- */
enum { TOPO_SURFACE = 0, TOPO_ALONG_Z = 1, TOPO_ALONG_X = 2, TOPO_NONE = 3 };

/* the branch the reference takes on nodeGlobal3DCounts (UpdateNode.c:
   196,248-251) */
static int topo_mode(const orc_context* c) {
    if (c->I > 2 && c->K > 2) return TOPO_SURFACE;
    if (c->I == 2 && c->K > 2) return TOPO_ALONG_Z;
    if (c->K == 2 && c->I > 2) return TOPO_ALONG_X;
    return TOPO_NONE;
}

/* a top node whose height the finite-difference cell defines */
static int top_interior(const orc_context* c, unsigned i, unsigned k) {
    const int inx = i > 0 && i < c->I - 1, inz = k > 0 && k < c->K - 1;
    switch (topo_mode(c)) {
    case TOPO_SURFACE: return inx && inz;
    case TOPO_ALONG_Z: return inz;
    case TOPO_ALONG_X: return inx;
    default: return 0;
    }
}

/* getMaxTopoGrad and its _alongZ/_alongX variants, UpdateNode.c:161-357
   (Cartesian branches) */
static void topo_max_grad(orc_context* c, unsigned n) {
    const unsigned i = n % c->I, j = (n / c->I) % c->J, k = n / (c->I * c->J);
    if (c->timeStep <= 1) return;
    if (j < c->J - 1) return;
    if (!top_interior(c, i, k)) return;
    const int mode = topo_mode(c);
    const double* X0 = c->coord[n];
    double g;
    if (mode == TOPO_SURFACE) {
        const double* xR = c->coord[n + 1];
        const double* xF = c->coord[n + c->I * c->J];
        const double dydx = (xR[1] - X0[1]) / (xR[0] - X0[0]);
        const double dydz = (xF[1] - X0[1]) / (xF[2] - X0[2]);
        g = sqrt(dydx * dydx + dydz * dydz);
    } else if (mode == TOPO_ALONG_Z) {
        const double* xF = c->coord[n + c->I * c->J];
        const double dydz = (xF[1] - X0[1]) / (xF[2] - X0[2]);
        g = sqrt(dydz * dydz);
    } else {
        const double* xR = c->coord[n + 1];
        const double dydx = (xR[1] - X0[1]) / (xR[0] - X0[0]);
        g = sqrt(dydx * dydx);
    }
    c->topoGradMax = c->topoGradMax > g ? c->topoGradMax : g;
}

/* Snac_UpdateNodeMomentum_DiffTopo (with diffTopo_alongZ/_alongX) and
   _UpdateInteriorTopo, UpdateNode.c:359-714 (Cartesian branches):
   dh = kappa dt lapl(y) from the five-point stencil (three-point along one
   direction), added to the node's height */
static void topo_diffuse(orc_context* c, unsigned n) {
    const unsigned i = n % c->I, j = (n / c->I) % c->J, k = n / (c->I * c->J);
    if (c->timeStep <= 1) return;
    if (j < c->J - 1 || c->topoGradMax < c->cfg.topoGradCriterion) return;
    if (!top_interior(c, i, k)) return;
    const int mode = topo_mode(c);
    const unsigned IJ = c->I * c->J;
    double* X0 = c->coord[n];
    double d2ydx2 = 0.0, d2ydz2 = 0.0;
    if (mode != TOPO_ALONG_Z) {
        const double* xR = c->coord[n + 1];
        const double* xL = c->coord[n - 1];
        const double dydx1 = (xR[1] - X0[1]) / (xR[0] - X0[0]);
        const double dydx2 = (X0[1] - xL[1]) / (X0[0] - xL[0]);
        d2ydx2 = (dydx1 - dydx2) / (0.5 * (xR[0] - xL[0]));
    }
    if (mode != TOPO_ALONG_X) {
        const double* xF = c->coord[n + IJ];
        const double* xB = c->coord[n - IJ];
        const double dydz1 = (xF[1] - X0[1]) / (xF[2] - X0[2]);
        const double dydz2 = (X0[1] - xB[1]) / (X0[2] - xB[2]);
        d2ydz2 = (dydz1 - dydz2) / (0.5 * (xF[2] - xB[2]));
    }
    const double lapl_topo = mode == TOPO_SURFACE ? d2ydx2 + d2ydz2
                           : (mode == TOPO_ALONG_Z ? d2ydz2 : d2ydx2);
    const double dh = c->cfg.topo_kappa * c->dt * lapl_topo;
    X0[1] += dh;
}

/* Snac_UpdateNodeMomentum_AdjustEdgeTopo, UpdateNode.c:716-857 (Cartesian
   branches): an edge node takes the height of its inner neighbour, a corner
   of the two-dimensional surface the mean of its two */
static void topo_adjust_edge(orc_context* c, unsigned n) {
    const unsigned i = n % c->I, j = (n / c->I) % c->J, k = n / (c->I * c->J);
    if (c->timeStep <= 1) return;
    if (j < c->J - 1 || c->topoGradMax < c->cfg.topoGradCriterion) return;
    const int mode = topo_mode(c);
    if (mode == TOPO_NONE) return;
    const unsigned IJ = c->I * c->J;
    const double *xN1 = NULL, *xN2 = NULL;
    if (mode != TOPO_ALONG_Z) {
        if (i == 0) xN1 = c->coord[n + 1];
        else if (i == c->I - 1) xN1 = c->coord[n - 1];
    }
    if (mode != TOPO_ALONG_X) {
        if (k == 0) xN2 = c->coord[n + IJ];
        else if (k == c->K - 1) xN2 = c->coord[n - IJ];
    }
    double* X0 = c->coord[n];
    if (xN1 && !xN2) X0[1] = xN1[1];
    else if (!xN1 && xN2) X0[1] = xN2[1];
    else if (xN1 && xN2) X0[1] = 0.5 * (xN1[1] + xN2[1]);
}

/* Snac_UpdateNodeMomentum, UpdateNode.c:60-72, and the topography hooks
   appended after it on Snac_EP_UpdateNodeMomentum (Context.c:578-592) */
int orc_update_node(orc_context* c, uint32_t n) {
    double* x = c->coord[n];
    const double* v = c->node[n].velocity;
    x[0] += c->dt * v[0];
    x[1] += c->dt * v[1];
    x[2] += c->dt * v[2];
    if (c->cfg.topo_kappa > 0) {
        topo_max_grad(c, n);
        topo_diffuse(c, n);
        topo_adjust_edge(c, n);
    }
    return SNH_OK;
}

/* _SnacRemesher_Remesh's copy-back and _SnacRemesher_UpdateElements
   (remesher_BI/Remesh.c:130-158, RemeshElements.c:133-151): the remeshed
   arrays replace the current ones, then every element is updated on the new
   mesh (old_volume = volume, new geometry, rotation with the current dt)
   and minLengthScale keeps the running minimum. */
int orc_remesh(orc_context* c, const snh_state* s) {
    int rc = orc_set_state(c, s);
    if (rc) return rc;
    for (unsigned e = 0; e < c->E; e++) {
        double m;
        orc_update_element(c, e, &m);
        if (m < c->minLengthScale) c->minLengthScale = m;
    }
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * SnacTemperature plugin: the energy entry points
 * ------------------------------------------------------------------------- */
static void apply_temperature_bc(orc_context* c) {
    for (unsigned n = 0; n < c->N; n++)
        if (c->tbc_mask[n]) c->temperature[n] = c->tbc_value[n];
}

/* _SnacTemperature_InitialConditions (temperature/VariableConditions.c:
   155-212): T0 = T (no restart), temperature BCs, zero heat flux.  The
   temperatureBCs composite of wall conditions: later entries win. */
int orc_set_thermal(orc_context* c, const snh_thermal* th) {
    if (!th || (th->nbcs && !th->bcs)) return fail("bad thermal description");
    if (!c->temperature0) {
        c->temperature0 = calloc(c->N, sizeof(double));
        c->heatFlux = calloc(c->E, sizeof *c->heatFlux);
        c->thermalStress = calloc(c->E, sizeof *c->thermalStress);
        c->tbc_mask = calloc(c->N, 1);
        c->tbc_value = calloc(c->N, sizeof(double));
        if (!c->temperature0 || !c->heatFlux || !c->thermalStress || !c->tbc_mask || !c->tbc_value)
            return fail("out of memory");
    }
    c->thermal = 1;
    c->thermal_stress = th->computeThermalStress != 0;
    if (th->temperature) memcpy(c->temperature, th->temperature, sizeof(double) * c->N);
    for (unsigned n = 0; n < c->N; n++) {
        c->tbc_mask[n] = 0;
        for (unsigned b = 0; b < th->nbcs; b++)
            if (node_on_wall(c, n, th->bcs[b].wall)) {
                c->tbc_mask[n] = 1;
                c->tbc_value[n] = th->bcs[b].value;
            }
    }
    memcpy(c->temperature0, c->temperature, sizeof(double) * c->N);
    apply_temperature_bc(c);
    memset(c->heatFlux, 0, sizeof(*c->heatFlux) * c->E);
    return SNH_OK;
}

/* Snac_HeatFlux (temperature/HeatFlux.c:63-155): per tet, k grad T from the
   stored geometry (node k pairs with surface (k+1)%4), and the thermal stress
   -alpha bulkm (mean T - mean T0). */
static void heat_flux(orc_context* c, unsigned e) {
    orc_element* el = &c->elem[e];
    const snh_material* m = &c->mat[el->material_I];
    const double bulkm = m->lambda + 2.0 * m->mu / 3.0;
    const double alpha = m->alpha;
    for (int t = 0; t < NTET; t++) {
        const orc_tet* tt = &el->tetra[t];
        const orc_surface* S = tt->surface;
        double T[4], T0[4];
        for (int k = 0; k < 4; k++) {
            unsigned nd = c->elementNodeTbl[e][TetraToNode[t][k]];
            T[k] = c->temperature[nd];
            T0[k] = c->temperature0[nd];
        }
        for (int d = 0; d < 3; d++)
            c->heatFlux[e][t][d] = 1.0 * m->thermal_conduct *
                                   (T[0] * S[1].normal[d] * S[1].area + T[1] * S[2].normal[d] * S[2].area +
                                    T[2] * S[3].normal[d] * S[3].area + T[3] * S[0].normal[d] * S[0].area) /
                                   3.0 / tt->volume;
        if (c->thermal_stress) {
            double temp = 0.25 * (T[0] + T[1] + T[2] + T[3]);
            double temp0 = 0.25 * (T0[0] + T0[1] + T0[2] + T0[3]);
            c->thermalStress[e][t] = -1.0 * alpha * bulkm * (temp - temp0);
        }
    }
}

/* SnacTemperature_HeatFluxBC (temperature/VariableConditions.c:227-353): on
   elements of the x and z walls, remove the normal component of -grad T
   (without the conductivity) from the flux; the wall normal is the mean of
   four tet surface normals, the last matching wall's. */
static void heat_flux_bc(orc_context* c, unsigned e) {
    static const unsigned wall_faces[4][4][2] = {
        {{0, 2}, {1, 2}, {5, 0}, {6, 0}},   /* i == 0 */
        {{2, 0}, {3, 0}, {7, 0}, {8, 0}},   /* i == EI-1 */
        {{0, 1}, {2, 2}, {6, 1}, {7, 1}},   /* k == 0 */
        {{1, 1}, {3, 2}, {5, 1}, {8, 1}}};  /* k == EK-1 */
    const unsigned i = e % c->EI, k = e / (c->EI * c->EJ);
    const int on[4] = {i == 0, i == c->EI - 1, k == 0, k == c->EK - 1};
    if (!(on[0] || on[1] || on[2] || on[3])) return;
    orc_element* el = &c->elem[e];
    double normal[3];
    for (int w = 0; w < 4; w++) {
        if (!on[w]) continue;
        for (int d = 0; d < 3; d++) {
            const unsigned(*f)[2] = wall_faces[w];
            normal[d] = 0.25 * (el->tetra[f[0][0]].surface[f[0][1]].normal[d] +
                                el->tetra[f[1][0]].surface[f[1][1]].normal[d] +
                                el->tetra[f[2][0]].surface[f[2][1]].normal[d] +
                                el->tetra[f[3][0]].surface[f[3][1]].normal[d]);
        }
    }
    for (int t = 0; t < NTET; t++) {
        const orc_tet* tt = &el->tetra[t];
        const orc_surface* S = tt->surface;
        double T[4], g[3], Fn;
        for (int kk = 0; kk < 4; kk++) T[kk] = c->temperature[c->elementNodeTbl[e][TetraToNode[t][kk]]];
        for (int d = 0; d < 3; d++)
            g[d] = -1.0 *
                   (T[0] * S[1].normal[d] * S[1].area + T[1] * S[2].normal[d] * S[2].area +
                    T[2] * S[3].normal[d] * S[3].area + T[3] * S[0].normal[d] * S[0].area) /
                   3.0 / tt->volume;
        Fn = 0.0;
        for (int d = 0; d < 3; d++) Fn += g[d] * normal[d];
        for (int d = 0; d < 3; d++) c->heatFlux[e][t][d] -= Fn * normal[d];
    }
}

/* SnacTemperature_LoopElements (temperature/HeatFlux.c:47-61) */
int orc_loop_elements_energy(orc_context* c) {
    if (!c->thermal) return fail("no temperature plugin (orc_set_thermal)");
    const long E = (long)c->E;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long e = 0; e < E; e++) heat_flux(c, (unsigned)e);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long e = 0; e < E; e++) heat_flux_bc(c, (unsigned)e);
    return SNH_OK;
}

/* Snac_Heat (temperature/Heat.c:64-132) for node n with sourceterm 0.
   Returns 0, or -1 when dt_thermal/dt_mech <= 1 (the firewall at :121-126). */
static int node_heat(orc_context* c, unsigned n) {
    double energy = 0.0, source = 0.0, lumpVolume = 0.0;
    const double sourceterm = 0.0;
    int bad = 0;
    for (unsigned ne = 0; ne < 8; ne++) {
        unsigned e = c->nodeElementTbl[n][ne];
        if (e >= c->E) continue;
        const orc_element* el = &c->elem[e];
        const snh_material* m = &c->mat[el->material_I];
        for (int idx = 0; idx < 5; idx++) {
            const unsigned t = NodeToTetra[ne][idx];
            const unsigned s = NodeToSurface[ne][idx];
            const orc_tet* tt = &el->tetra[t];
            const double* hf = c->heatFlux[e][t];
            const orc_surface* sf = &tt->surface[s];
            energy += (hf[0] * sf->normal[0] + hf[1] * sf->normal[1] + hf[2] * sf->normal[2]) * sf->area / 6.0;
            source += sourceterm * tt->volume / 4.0;
            lumpVolume +=
                m->heatCapacity *
                (((tt->density == 0.0)
                      ? (m->phsDensity * (1.0 - m->alpha * (tt->avgTemp - m->reftemp) + m->beta * el->hydroPressure))
                      : tt->density) *
                 tt->volume) /
                8.0;
        }
        double dt_thermal_to_mech = sqrt((m->lambda + 2.0 * m->mu) / m->phsDensity) * c->minLengthScale /
                                    (m->thermal_conduct / (m->heatCapacity * m->phsDensity));
        if (!(dt_thermal_to_mech > 1.0)) bad = 1;
    }
    c->temperature0[n] = c->temperature[n];
    c->temperature[n] += (energy + source) * -1.0 * c->dt / lumpVolume;
    return bad ? -1 : 0;
}

/* SnacTemperature_LoopNodes (temperature/Heat.c:47-62): nodal update, tet
   average temperature (UpdateAverageTemp_LoopElements, :135-160) before the
   BCs are re-applied, then the BCs. */
int orc_loop_nodes_energy(orc_context* c) {
    if (!c->thermal) return fail("no temperature plugin (orc_set_thermal)");
    const long N = (long)c->N, E = (long)c->E;
    int bad = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(min : bad)
    for (long n = 0; n < N; n++) {
        int r = node_heat(c, (unsigned)n);
        if (r < bad) bad = r;
    }
    if (bad) {
        snprintf(g_err, sizeof g_err, "timeStep=%u: dt_thermal / dt_mech <= 1 (Heat.c:121-126)", c->timeStep);
        return SNH_ENUMERIC;
    }
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long e = 0; e < E; e++) {
        orc_element* el = &c->elem[e];
        for (int t = 0; t < NTET; t++) {
            el->tetra[t].avgTemp = 0.0;
            for (int k = 0; k < 4; k++)
                el->tetra[t].avgTemp += c->temperature[c->elementNodeTbl[e][TetraToNode[t][k]]] / (1.0 * 4);
        }
    }
    apply_temperature_bc(c);
    return SNH_OK;
}

int orc_get_heat(orc_context* c, uint32_t e, uint32_t t, double heatFlux[3], double* thermalStress,
                 double* avgTemp, double* density) {
    if (!c->thermal) return fail("no temperature plugin (orc_set_thermal)");
    if (e >= c->E || t >= NTET) return fail("index out of range");
    for (int d = 0; d < 3; d++) heatFlux[d] = c->heatFlux[e][t][d];
    *thermalStress = c->thermalStress[e][t];
    *avgTemp = c->elem[e].tetra[t].avgTemp;
    *density = c->elem[e].tetra[t].density;
    return SNH_OK;
}

/* ---------------------------------------------------------------------------
 * loop drivers (Context.c:1102-1184) and Sync (:1208-1268)
 * ------------------------------------------------------------------------- */
int orc_calc_stresses(orc_context* c) {
    int rc = SNH_OK;
    const long E = (long)c->E;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(min : rc)
    for (long e = 0; e < E; e++) {
        orc_strain_rate(c, (unsigned)e);
        int r = orc_stress(c, (unsigned)e);
        if (r < rc) rc = r;
    }
    return rc;
}

int orc_loop_nodes(orc_context* c) {
    const long N = (long)c->N;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long n = 0; n < N; n++) {
        double mass, balance[3];
        orc_force(c, (unsigned)n, c->speedOfSound, &mass, &c->node[n].inertialMass, c->node[n].force, balance);
    }
    if (c->cfg.topo_kappa > 0) {
        /* the topography hooks read neighbours in node order: sequential */
        for (long n = 0; n < N; n++) {
            orc_update_node_preprocess(c, (unsigned)n, c->node[n].inertialMass, c->node[n].force);
            orc_update_node(c, (unsigned)n);
        }
        return SNH_OK;
    }
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (long n = 0; n < N; n++) {
        orc_update_node_preprocess(c, (unsigned)n, c->node[n].inertialMass, c->node[n].force);
        orc_update_node(c, (unsigned)n);
    }
    return SNH_OK;
}

int orc_loop_elements(orc_context* c) {
    const long E = (long)c->E;
    double gmin;
    orc_update_element(c, 0, &gmin);
    if (c->cfg.dtType == SNH_DT_WAVE) {
        /* the wave dt type updates speedOfSound inside the loop: keep it serial */
        for (long e = 1; e < E; e++) {
            double m;
            orc_update_element(c, (unsigned)e, &m);
            if (m < gmin) gmin = m;
        }
    } else {
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(min : gmin)
        for (long e = 1; e < E; e++) {
            double m;
            orc_update_element(c, (unsigned)e, &m);
            if (m < gmin) gmin = m;
        }
    }
    c->minLengthScale = gmin;
    return SNH_OK;
}

int orc_sync(orc_context* c) {
    double v, vmax = 0.0;
    const double vlb = F_3EM11;
    for (unsigned idx = 0; idx < c->bc_count; idx++) {
        const double* vel = c->node[c->bc_index[idx]].velocity;
        v = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
        if (v > vmax) vmax = v;
    }
    if (vmax < vlb) vmax = vlb;
    c->vmax = vmax;
    if (c->cfg.dtType == SNH_DT_DYNAMIC) {
        c->dt = c->minLengthScale * F_045 * c->strain_inert / vmax;
        c->speedOfSound = c->minLengthScale * 0.5 / c->dt;
    } else if (c->cfg.dtType == SNH_DT_WAVE) {
        c->dt = 0.5 * c->minLengthScale / c->speedOfSound;
    }
    return SNH_OK;
}

/* _SnacHydroStaticIC_IC (Snac/plugins/hydroStaticIC/VariableConditions.c:
   43-160), one rank: per element column, from the top down, an isotropic
   lithostatic stress.  Runs before TimeStepZero, when the tet densities are
   still zero. */
int orc_hydrostatic_ic(orc_context* c) {
    const unsigned elx = c->EI, ely = c->EJ, elz = c->EK;
    for (unsigned elK = 0; elK < elz; elK++)
        for (unsigned elI = 0; elI < elx; elI++) {
            double rogh = 0.0;
            for (int elJ = (int)ely - 1; elJ > -1; elJ--) {
                unsigned e = elI + (unsigned)elJ * elx + elK * elx * ely;
                orc_element* el = &c->elem[e];
                const snh_material* m = &c->mat[el->material_I];
                double densT = 0.0, dhE[4], dh, dPT, dP, P;
                for (int t = 0; t < NTET; t++) {
                    densT += ((el->tetra[t].density == 0.0)
                                  ? (m->phsDensity * (1.0 - m->alpha * (el->tetra[t].avgTemp - m->reftemp)))
                                  : (el->tetra[t].density)) /
                             NTET;
                    memset(el->tetra[t].strainRate, 0, sizeof el->tetra[t].strainRate);
                }
                dhE[0] = ecoord(c, e, 6)[1] - ecoord(c, e, 0)[1];
                dhE[1] = ecoord(c, e, 2)[1] - ecoord(c, e, 4)[1];
                dhE[2] = ecoord(c, e, 3)[1] - ecoord(c, e, 5)[1];
                dhE[3] = ecoord(c, e, 7)[1] - ecoord(c, e, 1)[1];
                dh = 0.25 * (dhE[0] + dhE[1] + dhE[2] + dhE[3]);
                el->stress = 0.0;
                el->strainRate = 0.0;
                el->hydroPressure = 0.0;
                dPT = densT * c->gravity * dh;
                dP = dPT * (1.0 - m->beta * rogh) / (1.0 + m->beta / 2.0 * dPT);
                P = rogh + 0.5 * dP;
                P = -1.0 * (P + 0.0);
                for (int t = 0; t < NTET; t++) {
                    el->tetra[t].stress[0][0] = P;
                    el->tetra[t].stress[1][1] = P;
                    el->tetra[t].stress[2][2] = P;
                }
                el->hydroPressure = P;
                rogh += dP;
            }
        }
    return SNH_OK;
}

/* Snac_Context_TimeStepZero, Context.c:978-1036 */
int orc_timestep_zero(orc_context* c) {
    c->topoGradMax = 0.0;   /* a run begins with it at zero (Context.c:759-764) */
    for (unsigned e = 0; e < c->E; e++) {
        orc_element* el = &c->elem[e];
        const snh_material* m = &c->mat[el->material_I];
        double sVolAvg = 0.0;
        el->stress = 0;
        for (int t = 0; t < NTET; t++) {
            const double(*st)[3] = el->tetra[t].stress;
            sVolAvg += st[1][1] * st[2][2] + st[2][2] * st[0][0] + st[0][0] * st[1][1];
            el->tetra[t].avgTemp = 0.0;
            el->tetra[t].density = m->phsDensity;
        }
        el->stress = 0.5 * sqrt(0.5 * fabs(sVolAvg / NTET));
    }
    c->timeStep = 0;
    orc_loop_elements(c);
    orc_sync(c);
    c->zero_done = 1;
    return SNH_OK;
}

int orc_step(orc_context* c, int nsteps) {
    if (!c->zero_done) {
        snprintf(g_err, sizeof g_err, "orc_step before orc_timestep_zero");
        return SNH_ESTATE;
    }
    for (int s = 0; s < nsteps; s++) {
        c->timeStep++;
        int rc;
        if (c->thermal) {   /* _Snac_Context_Solve, Context.c:1073-1074 */
            orc_loop_elements_energy(c);
            rc = orc_loop_nodes_energy(c);
            if (rc) return rc;
        }
        rc = orc_calc_stresses(c);
        if (rc) return rc;
        orc_loop_nodes(c);
        orc_loop_elements(c);
        orc_sync(c);
    }
    return SNH_OK;
}
