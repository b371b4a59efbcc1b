// inversion: drop-in for the reference's inversion.exe (src/inversion_main.cpp:3-32).
// Usage: inversion parameters.txt
#include <chrono>
#include <iostream>

#include "survey.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " parameters.txt\n";
        return 2;
    }
    try {
        fwi::Inversion inversion;
        inversion.parameters = argv[1];
        inversion.set_parameters();
        inversion.show_information();
        auto ti = std::chrono::system_clock::now();
        while (true) {
            inversion.compute_gradient();
            inversion.check_convergence();
            if (inversion.converged) break;
            inversion.optimization();
            inversion.update_model();
            inversion.save_checkpoint();
        }
        inversion.export_convergence();
        inversion.export_final_model();
        auto tf = std::chrono::system_clock::now();
        std::chrono::duration<double> e = tf - ti;
        std::cout << "\nRun time: " << e.count() << " s." << std::endl;
    } catch (const std::exception &ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
