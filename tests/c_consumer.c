/* A plain C99 caller of libspacetime_hip: includes the public header,
 * links the shared library, and exercises the host-only entry points a
 * foreign-function binding calls first (version string, camera block,
 * blackbody table, error codes).  Prints the results as text so the
 * Python test can compare them with the ctypes path.  No device work. */
#include <stdio.h>

#include "spacetime_hip.h"

int main(void) {
    const double pos[3] = {0.0, -30.0, 3.0}, target[3] = {0.0, 0.0, 0.0}, up[3] = {0.0, 0.0, 1.0};
    const double vel[3] = {0.0, 0.0, 0.0};
    double cam[49];
    float lut[16 * 3];
    int i, rc;

    printf("version %s\n", st_version());

    rc = st_camera_params(pos, target, up, 0.75, vel, 1.0, 0.9, 0.0, 1.0, 0, cam);
    printf("cam_rc %d\n", rc);
    printf("cam");
    for (i = 0; i < 49; ++i) printf(" %.17g", cam[i]);
    printf("\n");

    rc = st_blackbody_lut(6500.0, 500.0, 30000.0, 16, lut);
    printf("lut_rc %d\n", rc);
    printf("lut");
    for (i = 0; i < 16 * 3; ++i) printf(" %.9g", (double)lut[i]);
    printf("\n");

    /* spin above the extremal value: an error code and a message, no abort */
    rc = st_camera_params(pos, target, up, 0.75, vel, 1.0, 1.5, 0.0, 1.0, 0, cam);
    printf("bad_spin_rc %d message %s\n", rc, st_last_error());
    /* an empty blackbody table */
    rc = st_blackbody_lut(6500.0, 500.0, 30000.0, 0, lut);
    printf("empty_lut_rc %d message %s\n", rc, st_last_error());
    return 0;
}
