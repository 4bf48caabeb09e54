/* A plain C caller of the library, standing in for the Java FFM binding that this image
 * cannot compile: it drives the step loop through include/neofiz.h alone.
 *
 *   step_loop IN OUT
 *
 * IN (written by tests/test_gpu_c_client.py, little-endian):
 *   int32 N, E, A, hasJC, steps
 *   double r[N], z[N]; int32 conn[4E]; int32 axis[A]
 *   double material[6] (E, nu, rho, yield, Hiso, Hkin), jc[10], speed, friction
 * OUT: double ur[N], uz[N], vr[N], vz[N], then the solver time (nf_state.time), then the
 * queries temperature(0), plasticFrontRadius(), mostDamagedElement(), borePressure(),
 * meshZ(N - 1), the kinetic energy audit, and the discrete adjoint's gradient (6 values) of
 * the final axial displacement of node N - 1 (the run records checkpoints every 50 steps).
 *
 * The setup is the Taylor harness's (validate/TaylorImpactCase.java:225-249): axisymmetric,
 * reduced integration, finite strain, thickness 1, CFL 0.5, axis nodes pinned radially, a
 * rigid anvil at z = 0, a uniform downward velocity. */
#include <stdio.h>
#include <stdlib.h>

#include "neofiz.h"

static int fail(const char* what, int rc) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, nf_last_error());
    return 1;
}

static int read_all(FILE* f, void* p, size_t size, size_t count) {
    return fread(p, size, count, f) == count ? 0 : 1;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: step_loop IN OUT\n");
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return fail("open input", -1);
    int32_t head[5];
    if (read_all(in, head, sizeof(int32_t), 5)) return fail("read header", -1);
    const int32_t N = head[0], E = head[1], A = head[2], hasJC = head[3], steps = head[4];
    double* r = malloc(sizeof(double) * N);
    double* z = malloc(sizeof(double) * N);
    int32_t* conn = malloc(sizeof(int32_t) * 4 * (size_t)E);
    int32_t* axis = malloc(sizeof(int32_t) * (A > 0 ? A : 1));
    double m[6], jc[10], speed, friction;
    if (!r || !z || !conn || !axis) return fail("allocate", -1);
    if (read_all(in, r, sizeof(double), N) || read_all(in, z, sizeof(double), N) ||
        read_all(in, conn, sizeof(int32_t), 4 * (size_t)E) || read_all(in, axis, sizeof(int32_t), A) ||
        read_all(in, m, sizeof(double), 6) || read_all(in, jc, sizeof(double), 10) ||
        read_all(in, &speed, sizeof(double), 1) || read_all(in, &friction, sizeof(double), 1))
        return fail("read input", -1);
    fclose(in);

    nf_material mat = {m[0], m[1], m[2], m[3], m[4], m[5], hasJC, 0,
                       {jc[0], jc[1], jc[2], jc[3], jc[4], jc[5], jc[6], jc[7], jc[8], jc[9]}};
    nf_solver* s = NULL;
    int rc = nf_solver_create(&s, N, r, z, E, conn, 0, NULL, 1, &mat, NULL, NF_AXISYMMETRIC,
                              NF_REDUCED, NF_FINITE_STRAIN, 1.0, 0.5);
    if (rc) return fail("nf_solver_create", rc);
    for (int32_t k = 0; k < A; k++)
        if ((rc = nf_fix_radial(s, axis[k]))) return fail("nf_fix_radial", rc);
    if ((rc = nf_set_rigid_wall(s, 1, 0.0, friction))) return fail("nf_set_rigid_wall", rc);
    if ((rc = nf_set_uniform_velocity(s, 0.0, -speed))) return fail("nf_set_uniform_velocity", rc);
    if ((rc = nf_adjoint_enable(s, 50))) return fail("nf_adjoint_enable", rc);
    if ((rc = nf_run(s, steps))) return fail("nf_run", rc);

    double* buf = malloc(sizeof(double) * N);
    FILE* out = fopen(argv[2], "wb");
    if (!buf || !out) return fail("open output", -1);
    const int32_t fields[4] = {NF_NODE_UR, NF_NODE_UZ, NF_NODE_VR, NF_NODE_VZ};
    for (int k = 0; k < 4; k++) {
        if ((rc = nf_get_node_field(s, fields[k], buf))) return fail("nf_get_node_field", rc);
        if (fwrite(buf, sizeof(double), N, out) != (size_t)N) return fail("write output", -1);
    }
    nf_state st;
    if ((rc = nf_get_state(s, &st))) return fail("nf_get_state", rc);
    if (fwrite(&st.time, sizeof(double), 1, out) != 1) return fail("write output", -1);
    double q[6];
    int32_t worst = 0;
    nf_audits au;
    if ((rc = nf_temperature(s, 0, &q[0]))) return fail("nf_temperature", rc);
    if ((rc = nf_plastic_front_radius(s, &q[1]))) return fail("nf_plastic_front_radius", rc);
    if ((rc = nf_most_damaged_element(s, &worst))) return fail("nf_most_damaged_element", rc);
    q[2] = (double)worst;
    if ((rc = nf_bore_pressure(s, &q[3]))) return fail("nf_bore_pressure", rc);
    if ((rc = nf_mesh_z(s, N - 1, &q[4]))) return fail("nf_mesh_z", rc);
    if ((rc = nf_get_audits(s, &au))) return fail("nf_get_audits", rc);
    q[5] = au.kinetic_energy;
    if (fwrite(q, sizeof(double), 6, out) != 6) return fail("write output", -1);
    double* seed = calloc(4 * (size_t)N, sizeof(double));
    double grad[6];
    if (!seed) return fail("allocate", -1);
    seed[N + (N - 1)] = 1.0;  /* d / d uz of the last node */
    if ((rc = nf_adjoint_gradient(s, 1, seed, grad))) return fail("nf_adjoint_gradient", rc);
    if (fwrite(grad, sizeof(double), 6, out) != 6) return fail("write output", -1);
    free(seed);
    fclose(out);
    if ((rc = nf_solver_destroy(s))) return fail("nf_solver_destroy", rc);
    free(r);
    free(z);
    free(conn);
    free(axis);
    free(buf);
    return 0;
}
