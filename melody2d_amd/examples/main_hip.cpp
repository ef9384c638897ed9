// main_hip.cpp - the reference's driver (Code/main.cpp:83-238) with the step
// on the GPU, through the C-ABI of include/melody2d.h.
//   main_hip [number_save [saves [steps_per_save]]]
// reads STATIC_CONTROL.asc, STATIC_DATA.asc and DYNAMIC_<number_save>.asc of
// the working directory, then writes DYNAMIC_<number_save + k>.asc after
// every `steps_per_save` steps, k = 1 .. saves.
#include <cstdio>
#include <cstdlib>
#include "melody2d.h"

int main(int argc, char** argv) {
    const int number_save = argc > 1 ? atoi(argv[1]) : 0;
    const int saves = argc > 2 ? atoi(argv[2]) : 10;
    const int steps_per_save = argc > 3 ? atoi(argv[3]) : 1000;
    mel_ctx* ctx = nullptr;
    // Load_static + Load_dynamic + Code/main.cpp:103-126 (borders, BCs,
    // Update_proximity, Initialize_CZM, masses, Update_proximity, Initialize_CZM)
    if (mel_load(".", number_save, MEL_FP64, /*hip_device=*/0, &ctx) != MEL_OK) {
        fprintf(stderr, "load failed: %s\n", mel_last_error(nullptr));
        return 1;
    }
    // Monitoring after every iteration unless the model sets NO_MONITORING
    // (Code/main.cpp:172-176), written at every save (Code/main.cpp:221-222)
    const bool monitoring = mel_set_monitoring(ctx, 1) == MEL_OK;
    char name[64];
    int rc = 0;
    for (int save = number_save + 1; save <= number_save + saves; save++) {
        // steps_per_save x [Euler_step, or Solver_step for the adaptive schemes,
        // + contact update every Contact_update_period] (Code/main.cpp:148-196)
        mel_status st = mel_solver_steps(ctx, steps_per_save);
        if (st == MEL_OK) st = mel_synchronize(ctx);
        if (st != MEL_OK) {                                // MEL_E_NAN: Code/main.cpp:227-237
            fprintf(stderr, "step failed: %s\n", mel_last_error(ctx));
            rc = 2;
            break;
        }
        snprintf(name, sizeof name, "DYNAMIC_%05d.asc", save);
        if (mel_write_dynamic(ctx, name) != MEL_OK ||      // Write_dynamic, Code/IO.h:2351
            (monitoring && mel_write_monitoring(ctx, "MONITORING.asc") != MEL_OK)) {
            fprintf(stderr, "write failed: %s\n", mel_last_error(ctx));
            rc = 3;
            break;
        }
    }
    mel_info_t info;
    if (rc == 0 && mel_get_info(ctx, &info) == MEL_OK) printf("time %.17g iterations %d\n", info.time, info.iteration);
    mel_destroy(ctx);
    return rc;
}
