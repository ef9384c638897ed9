// main_facade.cpp - the same driver over the C++ façade of
// include/melody2d.hpp (Model / Body / Solver as Code/main.cpp uses them).
//   main_facade [number_save [saves [steps_per_save]]]
#include <cstdio>
#include <cstdlib>
#include <string>
#include "melody2d.hpp"

int main(int argc, char** argv) {
    const int number_save = argc > 1 ? atoi(argv[1]) : 0;
    const int saves = argc > 2 ? atoi(argv[2]) : 10;
    const int steps_per_save = argc > 3 ? atoi(argv[3]) : 1000;
    try {
        melody2d::Model model = melody2d::Model::load(".", number_save);
        melody2d::Solver solver;
        char name[64];
        for (int save = number_save + 1; save <= number_save + saves; save++) {
            // the model's scheme: Euler_step or Solver_step (Code/main.cpp:148-166)
            for (int k = 0; k < steps_per_save; k++) solver.Solver_step(model);
            model.synchronize();
            snprintf(name, sizeof name, "DYNAMIC_%05d.asc", save);
            model.write_dynamic(name);
        }
        double t, works[7], energy[3];
        model.scalars(t, works, energy);
        printf("time %.17g iterations %d bodies %d kinetic %.17g\n", t, model.Number_iteration(), model.Nb_bodies(),
               energy[0]);
    } catch (melody2d::Error& e) {
        fprintf(stderr, "melody2d error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
