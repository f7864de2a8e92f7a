/*
 * c_abi_consumer.c -- a plain C99 consumer of include/elsa_hip.h, linked against
 * libelsahip.so. It calls the host-only entry points (no GPU needed) the way an elsa-side
 * adapter or another C program would: the trajectory and pose helpers, the basis-function
 * tables, geometry creation and the error path. It prints one line per result as
 * "name v0 v1 ..." (%.9g), which tests/test_c_abi.py compares with the oracle.
 *
 * Built with: gcc -std=c99 -Wall -Wextra -pedantic -Werror -I include c_abi_consumer.c
 *             -L elsa_amd -lelsahip
 */
#include "elsa_hip.h"

#include <stdio.h>
#include <string.h>

static void print_floats(const char* name, const float* v, int n)
{
    int i;
    printf("%s", name);
    for (i = 0; i < n; ++i)
        printf(" %.9g", (double) v[i]);
    printf("\n");
}

static int check(int rc, const char* what)
{
    if (rc != EHP_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, ehp_last_error());
        return 1;
    }
    return 0;
}

int main(void)
{
    /* the 2D fan beam of the CPU reference workload: 256^2, 180 poses over 360 deg, det 384 */
    enum { NP2 = 180, NP3 = 4 };
    static float P2[NP2 * 6], Pi2[NP2 * 6], C2[NP2 * 2], E2[NP2 * 6], S2[NP2];
    static float P3[NP3 * 12], Pi3[NP3 * 12], C3[NP3 * 3], E3[NP3 * 12], S3[NP3];
    const int64_t vol2[2] = {256, 256}, det2[1] = {384};
    const int64_t vol3[3] = {64, 48, 40}, det3[2] = {96, 80};
    int64_t det_out[2] = {0, 0};
    float lut[100];
    double support = 0.0;
    float P[12], Pinv[12], C[3], ext[12], sdd = 0.0f;
    const float vsp[3] = {1.0f, 1.0f, 1.0f}, vorg[3] = {32.0f, 24.0f, 20.0f};
    const float dsp[2] = {1.0f, 1.0f}, dorg[2] = {48.0f, 40.0f}, ppo[2] = {0.0f, 0.0f}, cor[3] = {0, 0, 0};
    ehp_geometry* g = NULL;
    int rc;

    printf("version %s\n", ehp_version());

    if (check(ehp_circle_trajectory(2, NP2, vol2, 360.0f, 25600.0f, 1280.0f, det2, det_out, P2, Pi2, C2, E2, S2),
              "ehp_circle_trajectory 2D"))
        return 1;
    printf("det2 %lld\n", (long long) det_out[0]);
    print_floats("P2_0", P2, 6);
    print_floats("P2_45", P2 + 45 * 6, 6);
    print_floats("C2_179", C2 + 179 * 2, 2);
    print_floats("S2_0", S2, 1);

    if (check(ehp_circle_trajectory(3, NP3, vol3, 360.0f, 6400.0f, 320.0f, det3, det_out, P3, Pi3, C3, E3, S3),
              "ehp_circle_trajectory 3D"))
        return 1;
    print_floats("P3_1", P3 + 12, 12);
    print_floats("Pinv3_2", Pi3 + 24, 12);

    if (check(ehp_geometry_pose(3, 6400.0f, 320.0f, vsp, vorg, dsp, dorg, 0.5f, 0.25f, -0.125f, ppo, cor, P, Pinv,
                                C, ext, &sdd),
              "ehp_geometry_pose"))
        return 1;
    print_floats("pose_P", P, 12);
    print_floats("pose_C", C, 3);
    print_floats("pose_sdd", &sdd, 1);

    if (check(ehp_blob_lut(EHP_F32, 2.0, 10.83, 2, 100, lut, &support), "ehp_blob_lut"))
        return 1;
    print_floats("blob_lut", lut, 100);
    printf("blob_support %.9g\n", support);
    if (check(ehp_bspline_lut(EHP_F32, 3, 2, 100, lut, &support), "ehp_bspline_lut"))
        return 1;
    print_floats("bspline_lut", lut, 100);
    printf("bspline_support %.9g\n", support);

    /* geometry creation is host-only; an invalid dimension fails with a message */
    if (check(ehp_geometry_create(2, NP2, vol2, det2, P2, Pi2, C2, E2, S2, &g), "ehp_geometry_create"))
        return 1;
    ehp_geometry_destroy(g);
    g = NULL;
    rc = ehp_geometry_create(4, NP2, vol2, det2, P2, Pi2, C2, E2, S2, &g);
    printf("bad_dim %d %s %d\n", rc, g == NULL ? "null" : "set", (int) strlen(ehp_last_error()) > 0);
    rc = ehp_geometry_create(2, 0, vol2, det2, P2, Pi2, C2, E2, S2, &g);
    printf("no_poses %d\n", rc);
    printf("done\n");
    return 0;
}
