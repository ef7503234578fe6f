/* Host-side AddressSanitizer driver: exercises the C ABI's host code paths
 * (plain solver, graphs, snapshots, obstacle update, resize, batch admission
 * and retirement) on the GPU.  Built with -fsanitize=address for the host. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include "fvm/fvm.h"

#define CK(x) do { int rc_ = (x); if (rc_ != FVM_OK) { fprintf(stderr, "FAIL %s: %d %s\n", #x, rc_, fvm_last_error()); return 1; } } while (0)

static void viscous(fvm_params* p, int nx, int ny, int shape) {
    fvm_params_default(p);
    p->nx = nx; p->ny = ny;
    p->enable_viscosity = 1;
    p->obstacle_shape = shape;
    fvm_params_compute_derived(p);
}

int main(void) {
    setvbuf(stdout, NULL, _IONBF, 0);
    fvm_params p;
    viscous(&p, 512, 256, 0);
    fvm_solver* s = NULL;
    CK(fvm_create(0, &s));
    CK(fvm_initialize(s, &p));
    float dt = 0;
    CK(fvm_compute_stable_dt(s, &p, &dt));
    p.dt = dt;
    for (int n = 0; n < 5; n++) CK(fvm_step(s, &p));
    p.dt_mode = FVM_DT_DEVICE_CFL;
    CK(fvm_advance(s, &p, 37));
    CK(fvm_synchronize(s));
    printf("plain: %d steps\n", p.step);
    size_t n = (size_t)p.nx * p.ny;
    float* host = malloc(5 * n * sizeof(float));
    double* host64 = malloc(5 * n * sizeof(double));
    unsigned char* ct = malloc(n);
    CK(fvm_get_state(s, host));
    CK(fvm_get_state_f64(s, host64));
    CK(fvm_set_state(s, host));
    CK(fvm_get_cell_types(s, ct));
    CK(fvm_field_to_host(s, FVM_FIELD_MACH, host));
    double sums[4], f3[3];
    CK(fvm_conservation_sums(s, sums));
    CK(fvm_forces(s, f3));
    CK(fvm_save_snapshot(s, "/tmp/asan_snap.fvmsnap"));
    fvm_params q;
    CK(fvm_load_snapshot(s, "/tmp/asan_snap.fvmsnap", &q));
    p.obstacle_x += 0.2f;
    CK(fvm_update_obstacle(s, &p));
    CK(fvm_set_tile(s, 32, 16));
    CK(fvm_advance(s, &p, 20));
    CK(fvm_set_tile(s, 0, 0));
    CK(fvm_enable_timing(s, 1));
    CK(fvm_advance(s, &p, 4));
    CK(fvm_enable_timing(s, 0));
    CK(fvm_resize(s, 300, 200));
    viscous(&p, 300, 200, 3);
    CK(fvm_reset(s, &p));
    p.dt_mode = FVM_DT_DEVICE_CFL;
    CK(fvm_advance(s, &p, 18));
    CK(fvm_synchronize(s));
    CK(fvm_destroy(s));
    printf("plain: ok\n");

    fvm_batch* b = NULL;
    CK(fvm_batch_create(0, (size_t)4 << 30, 64, &b));
    int ids[40];
    for (int c = 0; c < 40; c++) {
        fvm_params r;
        viscous(&r, 64 + 17 * c, 48 + 9 * c, c % 6);
        r.enable_viscosity = c & 1;
        r.dt_mode = FVM_DT_DEVICE_CFL;
        CK(fvm_batch_add_case(b, &r, 30 + c, &ids[c]));
    }
    for (int round = 0; round < 6; round++) {
        CK(fvm_batch_advance(b, 16));
        for (int c = 0; c < 40; c++) {
            if (ids[c] < 0) continue;
            int32_t done = 0, fin = 0;
            double t = 0;
            CK(fvm_batch_case_status(b, ids[c], &done, &t, &fin));
            if (fin) {
                fvm_params r;
                viscous(&r, 64 + 17 * c, 48 + 9 * c, c % 6);
                float* st = malloc(5 * sizeof(float) * r.nx * r.ny);
                CK(fvm_batch_get_state(b, ids[c], st));
                free(st);
                if (c == 5) CK(fvm_batch_save_case(b, ids[c], "/tmp/asan_case.fvmsnap"));
                CK(fvm_batch_remove_case(b, ids[c]));
                ids[c] = -1;
            }
        }
    }
    int32_t id2 = -1;
    CK(fvm_batch_load_case(b, "/tmp/asan_case.fvmsnap", 50, &id2));
    CK(fvm_batch_advance(b, 10));
    int32_t act = 0;
    CK(fvm_batch_active_cases(b, &act));
    CK(fvm_batch_synchronize(b));
    CK(fvm_batch_destroy(b));
    printf("batch: ok (%d active at the end)\n", act);
    free(host); free(host64); free(ct);
    printf("driver: done\n");
    /* skip static destructors: the HIP runtime's teardown trips the
       sanitizer's device-allocator check at exit */
    _exit(0);
}
