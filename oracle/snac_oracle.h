/*
 * snac_oracle.h -- CPU oracle for SNAC's explicit time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm (geodynamics/snac), used by tests/ as the reference of the HIP path
 * and by bench.py as the timed CPU baseline.  The product (snac_amd/, the C ABI
 * in include/snac_hip.h) never links, loads or calls it.
 *
 * Each function cites the reference lines it restates.  It keeps the
 * reference's array-of-structures element layout, loop order, operand order,
 * float-literal constants, the int abs() truncation in the eigen-solver and the
 * unsymmetrised visco-plastic eigen-solve.  Build: gcc -O2 -ffp-contract=off
 * (the reference's x86-64 build has no FMA contraction).
 *
 * Parity pinning (what checks each part) is listed in the header of
 * snac_oracle.c.
 */
#ifndef SNAC_ORACLE_H
#define SNAC_ORACLE_H

#include "../include/snac_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_context orc_context;

/* per-tetrahedron record, as Snac_Element_Tetrahedra (Snac/libSnac/src/Element.h:51-61) */
typedef struct orc_tet_view {
    double old_volume, volume;
    double area[4];
    double normal[4][3];
    double strainRate[3][3];
    double strain[3][3];
    double stress[3][3];
    double density;
} orc_tet_view;

typedef struct orc_element_view {
    double volume, strainRate, stress, hydroPressure, aps;
} orc_element_view;

orc_context* orc_create(const snh_config* cfg);
void         orc_destroy(orc_context* c);
const char*  orc_last_error(void);

/* OpenMP threads for the element and node loops (1 = serial).  Results do
 * not depend on the thread count. */
void orc_set_threads(int n);
int  orc_get_max_threads(void);

/* 1 = apply velocity BCs by the reference's linear scan of the BC index table
 * (VariableCondition.c:459-463); 0 (default) = O(1) per-node table. */
void orc_set_bc_scan(orc_context* c, int faithful);
/* record, during the following steps, sign(v) sign(f) of every node component
   before damping ([N][3]) and the plastic branch of every tet ([E][10]:
   0 elastic, 1 shear, 2 tension); NULL stops recording */
void orc_record_decisions(orc_context* c, int8_t* damp, int8_t* branch);
/* per-node velocity-BC mask (bit d = component d constrained) and values [N][3] */
int orc_get_bc(orc_context* c, uint8_t* mask, double* value);

int orc_set_element_material(orc_context* c, const uint8_t* mat, size_t nelem);
int orc_set_element_field(orc_context* c, int field, const double* v, size_t nelem);
int orc_set_node_temperature(orc_context* c, const double* t, size_t nnode);
int orc_remesh(orc_context* c, const snh_state* s);                    /* remesher_BI/Remesh.c:130-158 */
int orc_set_thermal(orc_context* c, const snh_thermal* th);            /* temperature/VariableConditions.c:155-212 */
int orc_loop_elements_energy(orc_context* c);                          /* temperature/HeatFlux.c:47-155 */
int orc_loop_nodes_energy(orc_context* c);                             /* temperature/Heat.c:47-172 */
int orc_get_heat(orc_context* c, uint32_t e, uint32_t t, double heatFlux[3], double* thermalStress,
                 double* avgTemp, double* density);
int orc_set_state(orc_context* c, const snh_state* s);
int orc_get_state(orc_context* c, snh_state* s);
int orc_get_scalars(orc_context* c, snh_scalars* out);
int orc_set_dt(orc_context* c, double dt);
uint32_t orc_node_count(orc_context* c);
uint32_t orc_element_count(orc_context* c);

/* connectivity (HexaEL.c:456-571) */
int orc_element_nodes(orc_context* c, uint32_t* out /* [E][8] */);
int orc_node_elements(orc_context* c, uint32_t* out /* [N][8], sentinel = E */);

/* SnacHydroStaticIC initial stress (hydroStaticIC/VariableConditions.c:43-160) */
int orc_hydrostatic_ic(orc_context* c);

/* whole-step drivers */
int orc_timestep_zero(orc_context* c);  /* Context.c:978-1036 */
int orc_step(orc_context* c, int nsteps); /* AbstractContext.c:921-944 + Context.c:1063-1268 */

/* granular reference functions, for the reference's unit-test sequences */
int orc_update_element(orc_context* c, uint32_t e, double* minLen); /* UpdateElement.c:54-269 */
int orc_strain_rate(orc_context* c, uint32_t e);                     /* StrainRate.c:49-305 */
int orc_stress(orc_context* c, uint32_t e);                          /* Stress.c:53-172 */
int orc_test_stress(orc_context* c, uint32_t e);                     /* libSnac/tests/TestStress.c */
int orc_force(orc_context* c, uint32_t n, double speedOfSound,
              double* mass, double* inertialMass, double force[3], double balance[3]); /* Force.c:53-188 */
int orc_update_node_preprocess(orc_context* c, uint32_t n, double inertialMass, double force[3]); /* UpdateNode.c:74-158 */
int orc_update_node(orc_context* c, uint32_t n);                     /* UpdateNode.c:60-72 */
int orc_loop_elements(orc_context* c);                               /* Context.c:1156-1184 */
int orc_calc_stresses(orc_context* c);                               /* Context.c:1102-1117 */
int orc_loop_nodes(orc_context* c);                                  /* Context.c:1120-1154 */
int orc_sync(orc_context* c);                                        /* Context.c:1208-1268 */

int orc_get_tet(orc_context* c, uint32_t e, uint32_t t, orc_tet_view* out);
int orc_get_element(orc_context* c, uint32_t e, orc_element_view* out);
int orc_get_node_force(orc_context* c, uint32_t n, double force[3], double* inertialMass);

/* 1 (default) = tql2 uses the system libm hypot, as the reference; 0 = correctly rounded hypot */
void orc_set_glibc_hypot(int on);

/* eigen-decomposition of plastic_BI (Constitutive.c:240-510), for direct tests */
void orc_principal_stresses(const double stress[3][3], double d[3], double V[3][3]);
/* the x86 (int)abs(double) of the reference's tred2/tql2 */
int orc_ref_iabs(double x);

#ifdef __cplusplus
}
#endif
#endif
