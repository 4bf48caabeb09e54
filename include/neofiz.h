/*
 * neofiz.h -- C ABI of the MI355X explicit-dynamics step loop.
 *
 * The reference has no FFI: its boundary is the public Java API of
 * org.neofiz.solver.ExplicitSolver and its configuration objects. Each entry point below
 * replaces one member of that API; the citation names the Java member it stands for
 * (paths relative to src/main/java/org/neofiz/). A Java Panama FFM shim binds these symbols
 * one-for-one and presents the unchanged ExplicitSolver surface.
 *
 * Conventions
 *  - Every call returns an int status: NF_OK, or NF_E_ARGUMENT where the Java method throws
 *    IllegalArgumentException, NF_E_STATE where it throws IllegalStateException, NF_E_DEVICE
 *    for a HIP failure, NF_E_MEMORY when device memory runs out. nf_last_error() returns the
 *    message of the calling thread's last failure. No call aborts the process.
 *  - The library owns device memory. Arrays passed in are copied at the call; getters copy
 *    out into caller buffers (the Java accessors return live arrays; the shim re-reads them
 *    after each run call).
 *  - One handle is one HIP stream. Calls on one handle must be serialised by the caller;
 *    different handles may be driven from different threads.
 *  - Element connectivity is counter-clockwise in (r, z), 4 int32 node indices per element.
 */
#ifndef NEOFIZ_H
#define NEOFIZ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NF_ABI_VERSION 1

enum nf_status {
    NF_OK = 0,
    NF_E_ARGUMENT = -1, /* IllegalArgumentException */
    NF_E_STATE = -2,    /* IllegalStateException */
    NF_E_DEVICE = -3,   /* HIP runtime failure */
    NF_E_MEMORY = -4    /* device or host allocation failed */
};

/* core/Formulation.java:17-21 */
enum nf_formulation { NF_AXISYMMETRIC = 0, NF_PLANE_STRAIN = 1, NF_PLANE_STRESS = 2 };
/* solver/Integration.java:30-32 */
enum nf_integration { NF_FULL = 0, NF_REDUCED = 1 };
/* solver/Kinematics.java:20, :35 */
enum nf_kinematics { NF_SMALL_STRAIN = 0, NF_FINITE_STRAIN = 1 };

/* core/JohnsonCook.java:68-71 -- the record's components, in order. */
typedef struct nf_johnson_cook {
    double a, b, n, c, m;
    double reference_strain_rate, room_temperature, melt_temperature;
    double specific_heat, taylor_quinney;
} nf_johnson_cook;

/* core/Material.java:17-19 -- the record's components. yield_stress may be +infinity
 * (the elastic material, core/Material.java:54-56). */
typedef struct nf_material {
    double youngs_modulus, poisson_ratio, density;
    double yield_stress, isotropic_hardening, kinematic_hardening;
    int32_t has_johnson_cook;
    int32_t reserved;
    nf_johnson_cook johnson_cook;
} nf_material;

typedef struct nf_solver nf_solver;

/* Message of the calling thread's last failure, or "" after a success. */
const char* nf_last_error(void);

/* Validates a material as the Material and JohnsonCook compact constructors do
 * (core/Material.java:21-37, core/JohnsonCook.java:73-89). */
int nf_material_check(const nf_material* material);

/*
 * new ExplicitSolver(mesh, materials, formulation, integration, kinematics, thickness, cfl)
 * solver/ExplicitSolver.java:264-339, with new QuadMesh(r, z, conn, pressureEdges)
 * (mesh/QuadMesh.java:37-54) and Materials.uniform/of (core/Materials.java:36-53).
 * material_of is NULL for a uniform mesh (material_count must then be 1), else one index
 * per element into materials.
 */
int nf_solver_create(nf_solver** out, int32_t node_count, const double* r, const double* z,
                     int32_t element_count, const int32_t* conn,
                     int32_t pressure_edge_count, const int32_t* pressure_edges,
                     int32_t material_count, const nf_material* materials,
                     const int32_t* material_of, int32_t formulation, int32_t integration,
                     int32_t kinematics, double thickness, double cfl_safety);

/* ExplicitSolver.shutdown() (solver/ExplicitSolver.java:395-400) plus release of the handle.
 * NULL is accepted; calling it once per handle is the caller's duty. */
int nf_solver_destroy(nf_solver* s);

/* setThreads(n) solver/ExplicitSolver.java:386-392. n >= 1; the answer never depends on it. */
int nf_set_threads(nf_solver* s, int32_t n);
/* fixRadial / fixAxial / fixAllAxial :408-425 */
int nf_fix_radial(nf_solver* s, int32_t node);
int nf_fix_axial(nf_solver* s, int32_t node);
int nf_fix_all_axial(nf_solver* s);
/* setPressureRamp :432-435 */
int nf_set_pressure_ramp(nf_solver* s, double pascals, double ramp_seconds);
/* setRelaxationDamping :450-455 */
int nf_set_relaxation_damping(nf_solver* s, double fraction_of_critical,
                              double angular_frequency);
/* setHourglassCoefficient :466-469 */
int nf_set_hourglass_coefficient(nf_solver* s, double q);
/* setRigidWall(RigidWall.atZ(z, mu)) :475-477, solver/RigidWall.java:109-136;
 * enabled = 0 is setRigidWall(null). friction may be +infinity (welded). */
int nf_set_rigid_wall(nf_solver* s, int32_t enabled, double plane_z, double friction);
/* setContact(new Contact(mesh).withFriction(mu).withDamping(zeta).withStiffnessScale(k))
 * :487-489, solver/Contact.java:273-364; enabled = 0 is setContact(null). */
int nf_set_contact(nf_solver* s, int32_t enabled, double friction, double damping,
                   double stiffness_scale);
/* Test hook with no counterpart in the reference: lowers two capacities of the contact
 * search so that its rare paths can be exercised. capture_capacity caps the entries of the
 * capture grid (0 sends every assembly down the reference's own nine-cell walk, the path
 * taken when the grid overflows; the pairs and forces are the same); max_terms caps the pair
 * terms one node may take in one step (at most 32; a node over the cap makes the run call
 * fail with NF_E_STATE after its steps). -1 restores either default. NF_E_STATE without
 * contact. */
int nf_contact_set_limits(nf_solver* s, int32_t capture_capacity, int32_t max_terms);
/* setGravity :509-512 */
int nf_set_gravity(nf_solver* s, double acceleration_r, double acceleration_z);
/* setDamage(double[] fractureEnergy, double[] failureStrain) :729-791, one per element. */
int nf_set_damage(nf_solver* s, const double* fracture_energy, const double* failure_strain);
/* setErosion :549-561 */
int nf_set_erosion(nf_solver* s, double damage_threshold);
/* setMassScaling :601-619 */
int nf_set_mass_scaling(nf_solver* s, double target_step);
/* setVelocity / setUniformVelocity :799-808 */
int nf_set_velocity(nf_solver* s, int32_t node, double velocity_r, double velocity_z);
int nf_set_uniform_velocity(nf_solver* s, double velocity_r, double velocity_z);

/* run(steps) :936-938 */
int nf_run(nf_solver* s, int64_t steps);
/* while (time() < end_time) step(), validate/TaylorImpactCase.java:256-259, stopping early
 * after max_steps (<= 0 means no limit). steps_taken may be NULL. */
int nf_run_until(nf_solver* s, double end_time, int64_t max_steps, int64_t* steps_taken);
/* advance() :1069-1075 */
int nf_advance(nf_solver* s);

/* Node fields, node_count doubles each. radialDisplacement :842, axialDisplacement :846,
 * velocityR/Z :905-912, forceR/Z :915-922; mass and the reference coordinates are the
 * solver's private arrays (:107, mesh/QuadMesh.java:21-23); ON_WALL, FIXED_R/Z are 0 or 1. */
enum nf_node_field {
    NF_NODE_UR = 0, NF_NODE_UZ, NF_NODE_VR, NF_NODE_VZ, NF_NODE_FR, NF_NODE_FZ,
    NF_NODE_MASS, NF_NODE_R, NF_NODE_Z, NF_NODE_ON_WALL, NF_NODE_FIXED_R, NF_NODE_FIXED_Z,
    NF_NODE_FIELD_COUNT
};
int nf_get_node_field(nf_solver* s, int32_t field, double* out);
/* Writes UR, UZ, VR or VZ (the mutation a Java caller makes through a live array). */
int nf_put_node_field(nf_solver* s, int32_t field, const double* in);

/* Gauss-point fields: stress() :930 (4 per point {rr, zz, tt, rz}), back stress (4 per
 * point), plasticStrain() :925, plasticWorkDensity() :1712, and the damage state
 * (solver/Damage.java:188-192: strength, softening, band width). The rest serve the damage()
 * object (:794) and temperature(int) (:1832) in bulk:
 *   TEMPERATURE            temperature(p) at every point (JohnsonCook.temperature,
 *                          core/JohnsonCook.java:218-221); NaN at a point whose material has
 *                          no Johnson-Cook law (where temperature(p) throws)
 *   DAMAGE                 Damage.at(p, plasticStrain()[p]) (solver/Damage.java:342-347)
 *   DAMAGE_FAILURE_STRAIN  the point's failure strain, DAMAGE_FRACTURE_ENERGY
 *                          Damage.fractureEnergy(p) (solver/Damage.java:266-268)
 * The damage fields read 0 when no damage model is set (damage() returns null). */
enum nf_point_field {
    NF_POINT_STRESS = 0, NF_POINT_BACK, NF_POINT_PLASTIC_STRAIN, NF_POINT_PLASTIC_WORK,
    NF_POINT_DAMAGE_STRENGTH, NF_POINT_DAMAGE_SOFTENING, NF_POINT_DAMAGE_WIDTH,
    NF_POINT_TEMPERATURE, NF_POINT_DAMAGE, NF_POINT_DAMAGE_FAILURE_STRAIN,
    NF_POINT_DAMAGE_FRACTURE_ENERGY,
    NF_POINT_FIELD_COUNT
};
int nf_get_point_field(nf_solver* s, int32_t field, double* out);

/* Element fields: hourglass force hgR/hgZ :134-135, stiffness :137, eroded flag isEroded :627
 * (0 or 1). */
enum nf_element_field {
    NF_ELEMENT_HG_R = 0, NF_ELEMENT_HG_Z, NF_ELEMENT_HG_STIFFNESS, NF_ELEMENT_ERODED,
    NF_ELEMENT_FIELD_COUNT
};
int nf_get_element_field(nf_solver* s, int32_t field, double* out);

/* Scalar state: time() :834, timestep() :830, steps() :838, massScale() :622,
 * erodedElements/erodedEnergy/invertedElements :632-652, RigidWall audits
 * (solver/RigidWall.java:285-315) and Contact audits (solver/Contact.java:371-404). */
typedef struct nf_state {
    double time, dt, dt_previous, dt0, mass_scale;
    int64_t steps;
    int32_t node_count, element_count, points_per_element, eroded_elements;
    int32_t inverted_elements, wall_contact_count, wall_stuck_count, contact_pairs;
    double eroded_energy;
    double wall_impulse, wall_energy_loss, wall_friction_dissipation;
    double contact_energy, contact_dissipation, contact_max_penetration;
    int64_t contact_escaped_nodes;
} nf_state;
int nf_get_state(nf_solver* s, nf_state* out);

/* Single-value queries of ExplicitSolver (solver/ExplicitSolver.java):
 *   nf_temperature          temperature(int point) :1832-1840; NF_E_STATE when the point's
 *                           material has no Johnson-Cook law, NF_E_ARGUMENT out of range
 *   nf_plastic_front_radius plasticFrontRadius() :1867-1877 (largest reference-mesh corner
 *                           radius of an element with a yielded point)
 *   nf_most_damaged_element mostDamagedElement() :1778-1794 (-1 without damage or before any
 *                           softening)
 *   nf_bore_pressure        borePressure() :826-828 (the ramped pressure at the current time)
 *   nf_mesh_z               meshZ(int node) :1928-1930 */
int nf_temperature(nf_solver* s, int32_t point, double* out);
int nf_plastic_front_radius(nf_solver* s, double* out);
int nf_most_damaged_element(nf_solver* s, int32_t* out);
int nf_bore_pressure(nf_solver* s, double* out);
int nf_mesh_z(nf_solver* s, int32_t node, double* out);

/* setDamage(double fractureEnergy, double[] failureStrain) :713-717: one fracture energy
 * for every element, then as nf_set_damage. */
int nf_set_damage_uniform(nf_solver* s, double fracture_energy, const double* failure_strain);

/* Audits, solver/ExplicitSolver.java:875-886, :1610-1860. max_temperature is NaN when no
 * material has a Johnson-Cook law (temperature() throws there, :1834-1837). */
typedef struct nf_audits {
    double axial_momentum, radial_momentum, kinetic_energy, centred_kinetic_energy;
    double strain_energy, plastic_dissipation, hourglass_energy, fracture_dissipation;
    double max_plastic_strain, yielded_fraction, max_temperature, max_damage;
    double damaged_fraction, failed_fraction;
} nf_audits;
int nf_get_audits(nf_solver* s, nf_audits* out);

/* Contact queries (solver/Contact.java:371-404 and the pair set the search chose, :446-462).
 * nf_contact_pairs: for every surface node, in ascending node order, that resolved a pair
 * in the last assembly, {node, segment end a, segment end b}; nf_contact_surface: the
 * surface segments {a, b} in surface order (element by element, mesh/QuadMesh.java:188-226,
 * over the elements the contact has not forgotten). count receives the total; at most
 * capacity entries are written. Both report nothing when contact is off. */
int nf_contact_pairs(nf_solver* s, int32_t capacity, int32_t* out, int32_t* count);
int nf_contact_surface(nf_solver* s, int32_t capacity, int32_t* out, int32_t* count);

/* Parameter sensitivities (no counterpart in the reference): the exact derivative of the
 * discrete step loop with respect to {anvil friction, Johnson-Cook A, B, n, C, m}, carried
 * forward alongside the state in those six directions. nf_enable_sensitivity switches it on
 * before the first step, for reduced integration of one Johnson-Cook material without
 * kinematic hardening, damage, erosion, contact or pressure edges (else NF_E_STATE); the
 * trajectory stays bit-identical to a run without it. nf_get_node_sensitivity writes
 * d(field)/d(parameter) per node for UR, UZ, VR, VZ, FR, FZ; nf_get_step_sensitivity writes
 * the derivatives of the current and the previous time step (2 x 6 doubles). advance() is
 * refused while sensitivities are on; the wall audits are not accumulated. */
int nf_enable_sensitivity(nf_solver* s);
int nf_get_node_sensitivity(nf_solver* s, int32_t field, int32_t parameter, double* out);
int nf_get_step_sensitivity(nf_solver* s, double* out);

/* Discrete adjoint (no counterpart in the reference): the reverse-mode gradient of
 * objectives of the final nodal state with respect to {anvil friction, Johnson-Cook A, B, n,
 * C, m}, for the same configurations as the sensitivities above. nf_adjoint_enable (before
 * the first step) makes every later run call record a checkpoint of the whole state every
 * checkpoint_interval steps (1 .. 4096); the trajectory is the plain one, bit for bit.
 * nf_adjoint_gradient then runs the reverse sweep over every recorded step: for each
 * checkpoint, from the last, it reruns its steps recording what each step read, and runs
 * them backwards. seeds holds, per objective, its derivatives with respect to the final
 * UR, UZ, VR and VZ ([objective_count][4][node_count]); gradient receives [objective_count][6].
 * At most two objectives per sweep. The solver is left in its final state. advance() is
 * refused, and contact, damage and erosion cannot be set, while the adjoint is enabled.
 * nf_adjoint_info writes {recorded steps, checkpoints, element passes differentiated through
 * a Newton solve stopped on its iteration cap in the last sweep, checkpoint interval}. */
int nf_adjoint_enable(nf_solver* s, int32_t checkpoint_interval);
int nf_adjoint_gradient(nf_solver* s, int32_t objective_count, const double* seeds,
                        double* gradient);
int nf_adjoint_info(nf_solver* s, int64_t* out);

/* Batched ensemble (no single counterpart in the reference, which runs an ensemble as one
 * solver per thread: M0Cylinder.java:870-907, sweep/Sweep.java:77-106). A batch is made from
 * solvers already set up through the calls above (mesh, material, wall, velocities ...) and
 * steps them all together on the GPU, each with its own time step and end time; every
 * instance's state stays bit-identical to what a solo run would give. nf_batch_scatter
 * copies each instance's state back into its solver, whose getters then read it. A batch
 * refuses solvers with contact, damage, erosion, pressure edges or several materials
 * (NF_E_STATE), and mixed integration rules (NF_E_ARGUMENT); the wall audits of the
 * solvers are not advanced by a batch. The solvers must outlive the batch.
 * nf_batch_set_tile sets the threads per tile (64, 128 or 256; instance -1 = all) that an
 * instance's elements and nodes are cut into; nf_batch_run_until runs every instance while
 * its time is below end_time, at most max_steps steps each when max_steps > 0. */
typedef struct nf_batch nf_batch;
int nf_batch_create(nf_batch** out, int32_t count, nf_solver* const* solvers);
int nf_batch_destroy(nf_batch* b);
int nf_batch_set_tile(nf_batch* b, int32_t instance, int32_t threads);
int nf_batch_get_tile(nf_batch* b, int32_t instance, int32_t* threads);
int nf_batch_run_until(nf_batch* b, double end_time, int64_t max_steps);
int nf_batch_scatter(nf_batch* b);

/* ---- Film frames (render/Film.java) ----
 *
 * Film.capture(solver) (render/Film.java:159-189) for the element quantities the reference's
 * scenes colour by, computed on the device from the solver's state in one launch:
 *   PLASTIC_STRAIN  elementPlasticStrain (solver/ExplicitSolver.java:1897-1901)
 *   DAMAGE          elementDamage (:1764-1772)
 *   SPEED           elementSpeed (:895-902)
 *   MISES           von Mises of centroidStress (:1886-1894; Impact.java:229-232)
 *   STRESS_RR..RZ   centroidStress components
 *   CENTROID_Z      centroidZ in the current configuration (:1918-1925)
 *   ALIVE           1 - isEroded (:627-629), the usual aliveIf
 * r and z receive the deformed node positions (mesh + displacement, Film.java:161-166);
 * values receives field_count x element_count values, field-major. Arbitrary caller-defined
 * quantities (the Java lambdas) are host-side and not part of this call. */
enum nf_film_field {
    NF_FILM_PLASTIC_STRAIN = 0,
    NF_FILM_DAMAGE = 1,
    NF_FILM_SPEED = 2,
    NF_FILM_MISES = 3,
    NF_FILM_STRESS_RR = 4,
    NF_FILM_STRESS_ZZ = 5,
    NF_FILM_STRESS_TT = 6,
    NF_FILM_STRESS_RZ = 7,
    NF_FILM_CENTROID_Z = 8,
    NF_FILM_ALIVE = 9
};
int nf_film_capture(nf_solver* s, int32_t field_count, const int32_t* fields, double* r, double* z,
                    double* values);

/* ---- Defect-field seeding (core/DefectField.java, core/Weibull.java, core/Normal.java) ----
 *
 * new DefectField(new Weibull(scale, modulus), referenceVolume, correlationLength, seed)
 * (core/DefectField.java:123-142, core/Weibull.java:53-69): a spatially correlated standard
 * normal field (stateless SplitMix64 white noise on a grid of correlationLength/3, smoothed by
 * a separable Gaussian) mapped through the normal cdf onto the Weibull that an element's
 * effective volume earns it. Every value is a function of (seed, position) alone, so two
 * meshes of one body see the same field bit for bit. The handle owns a HIP stream and the
 * scratch it sizes on demand (the white-noise grid over the mesh's bounding box). */
typedef struct nf_defect_params {
    double scale;              /* Weibull scale at the reference volume */
    double modulus;            /* Weibull modulus m */
    double reference_volume;   /* V0, m^3 */
    double correlation_length; /* l, m */
    int64_t seed;
} nf_defect_params;

typedef struct nf_defect nf_defect;

/* The constructor's checks, in Java's order: Weibull scale and modulus positive
 * (core/Weibull.java:62-69), then referenceVolume and correlationLength positive
 * (core/DefectField.java:124-131); NaN fails each. A non-finite correlationLength, which the
 * reference accepts and degenerates on, is refused. */
int nf_defect_create(nf_defect** out, const nf_defect_params* params);
int nf_defect_destroy(nf_defect* f);

/* Effective-volume rule of a sampled element. RING is DefectField.effectiveVolume
 * (:204-208), the rule of sampleOnto (:239-248): l^3 max(1, A/l^2) max(1, 2 pi rc/l).
 * SLAB has no counterpart in the reference; it is the plane-strain analogue, the same
 * in-plane count times the number of correlation lengths through the depth:
 * l^3 max(1, A/l^2) max(1, depth/l), passed through failureStrainForVolume (:226-228). */
enum nf_defect_volume { NF_DEFECT_RING = 0, NF_DEFECT_SLAB = 1 };

/* Flag: hash every white-noise tap instead of reading the memoised grid (the reference's
 * path when the grid would exceed its cache, :365-366). The values are identical. */
#define NF_DEFECT_UNCACHED 1

/* sampleOnto(QuadMesh) :239-248, for seed_count fields that differ only in their seed
 * (seeds NULL: the handle's own seed, seed_count 1). Element centroids are area centroids
 * and areas signed areas of the reference configuration (mesh/QuadMesh.java:78-118). out
 * receives seed_count x element_count failure strains, seed-major. Host arrays in and out. */
int nf_defect_sample_onto(nf_defect* f, int32_t seed_count, const int64_t* seeds,
                          int32_t node_count, const double* r, const double* z,
                          int32_t element_count, const int32_t* conn, int32_t volume_rule,
                          double depth, int32_t flags, double* out);

/* The same with the mesh and the output already in device memory (r, z, conn, out device
 * pointers; seeds a host array). Runs on the handle's stream and returns when it is done. */
int nf_defect_sample_onto_device(nf_defect* f, int32_t seed_count, const int64_t* seeds,
                                 int32_t node_count, const double* r, const double* z,
                                 int32_t element_count, const int32_t* conn,
                                 int32_t volume_rule, double depth, int32_t flags, double* out);

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * inside a captured stream, tight serving loops, harnesses that forbid hipMalloc). The same
 * as nf_defect_sample_onto_device, bit for bit, but the call allocates nothing: its bounding
 * box, status word, seed hashes and white-noise grid live in the library's static 16 MiB
 * device array (one per GPU, held until the process ends), not in the handle's grown
 * buffers. seed_count is at most 16384 (else NF_E_ARGUMENT); a box whose grid exceeds the
 * array even for one seed hashes every tap instead (the same values). Calls that overlap on
 * one GPU share that array and race: the caller serialises them. */
int nf_defect_sample_onto_static(nf_defect* f, int32_t seed_count, const int64_t* seeds,
                                 int32_t node_count, const double* r, const double* z,
                                 int32_t element_count, const int32_t* conn,
                                 int32_t volume_rule, double depth, int32_t flags, double* out);

/* Point evaluations at n points (host arrays):
 *   GAUSSIAN          gaussianAt(r, z) :172-174                      (w unused)
 *   STRAIN_AREA       failureStrainAt(r, z, crossSectionArea = w) :217-219
 *   STRAIN_VOLUME     failureStrainForVolume(r, z, volume = w) :226-228
 *   EFFECTIVE_VOLUME  effectiveVolume(centroidRadius = r, crossSectionArea = w) :204-208
 *   TRANSFORM         transform(g = r, volume = w) :261-271 (private in the reference)
 *   WHITE             white(i = r, j = z) :325-332 (private in the reference; r and z must
 *                     be integers in int32 range)
 * A volume that is not positive, or one whose weakest-link scale underflows, fails with
 * NF_E_ARGUMENT as Weibull.forVolume / minimumOf / the Weibull record do. */
enum nf_defect_function {
    NF_DEFECT_GAUSSIAN = 0,
    NF_DEFECT_STRAIN_AREA = 1,
    NF_DEFECT_STRAIN_VOLUME = 2,
    NF_DEFECT_EFFECTIVE_VOLUME = 3,
    NF_DEFECT_TRANSFORM = 4,
    NF_DEFECT_WHITE = 5
};
int nf_defect_points(nf_defect* f, int32_t function, int32_t n, const double* r,
                     const double* z, const double* w, double* out);

/* Device milliseconds of the last sampling call's kernels, measured with HIP events on the
 * handle's stream: {bounding box, white-noise grid, smoothing and transform}. */
int nf_defect_timing(nf_defect* f, double* ms);

/* Normal.erfc (0), erf (1), cdf (2), survival (3) (core/Normal.java:39-89), Cody's rational
 * approximations, evaluated on the device at n points. */
int nf_normal_eval(int32_t function, int32_t n, const double* x, double* out);

/* Per-kernel timing of the step loop (no counterpart in the reference, which times whole
 * runs with System.nanoTime, validate/TaylorImpactCase.java:254-260). When enabled, each
 * launch is bracketed by HIP events on the handle's stream; nf_profile_read returns summed
 * milliseconds and launch counts for {node pass, element pass, contact launches, fused step
 * kernel} (arrays of four) and resets them. */
int nf_profile(nf_solver* s, int32_t enable);
int nf_profile_read(nf_solver* s, double* ms, int64_t* counts);

/* J2.update (solver/J2.java:136-296) at `paths` independent material points of one material:
 * increments holds, per path and step, {dRR, dZZ, dTT, dGamRZ (engineering), strainDt};
 * damage is NULL or, per path, {failureStrain, fractureEnergy, area, uR, uZ, vR, vZ} (the
 * Damage point data of solver/Damage.java:222-263 and ExplicitSolver.setDamage :767-789);
 * out receives per path and step {sig rr, zz, tt, rz, back rr, zz, tt, rz, epsP, plastic
 * work, dGamma, damage strength, softening, band width}. formulation selects the plane-stress
 * lambda of :291-293. Points start from zero stress. */
int nf_point_path(const nf_material* material, int32_t formulation, int32_t paths,
                  int32_t steps, const double* increments, const double* damage, double* out);

/* Device-to-device copy bandwidth probe (no counterpart in the reference): copies `bytes`
 * reps times with a plain 16-byte-per-thread kernel; reports GB/s (read + write) and ms per
 * copy. It is the measured stream-copy peak the step loop's HBM fraction is also quoted
 * against, and the known-size kernel that hardware counters are calibrated with. */
int nf_copy_bandwidth(int64_t bytes, int32_t reps, double* gbps, double* ms_per_copy);

/* Device self-check: evaluates the GPU's log (0), exp (1), pow (2), atan2 (3), sin (4),
 * cos (5), hypot (6), sqrt (7), division (8), log1p (9), or the Newton loop's pow (10) and
 * log (11) with special inputs out of line, and the derivative pass's pow (12) with the log
 * it shares (13) (the same results as 2, 0, 2 and 0), the Newton loop's pow(x, y) (14) and
 * log(y) (15) evaluated side by side, and the yield check's pow(x, y) (16) and pow(y, x)
 * (17) evaluated side by side (the same results as pow and log) on n inputs, the hardware's
 * single-precision log2 (18), exp2 (19) and reciprocal (20) of the inputs rounded to single
 * precision, and the Johnson-Cook multiplier solve (solver/J2.java:332-375) with its leading
 * bisections proven rather than evaluated (21), with every iteration evaluated (22; the same
 * results as 21), and the number of bisections 21 proved (23); for 21-23, x holds five
 * values per solve (norm, stiffness, plastic strain, thermal factor, dt) and y the five
 * material values A, B, n, C and the reference strain rate; and the Newton loop's pow(x, y)
 * (24) and log(y) (25) for an exponent 0 < y < 1, without exp's range selects (the same
 * results as pow and log). These are the
 * functions the kernels use in place of java.lang.Math's (core/JohnsonCook.java:139, :182,
 * :196; solver/Damage.java:315-322; core/Weibull.java:118, :130; core/DefectField.java:331),
 * evaluated where the kernels run. */
int nf_math_eval(int32_t fn, const double* x, const double* y, double* out, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* NEOFIZ_H */
