// modeling: drop-in for the reference's modeling.exe (src/modeling_main.cpp:3-29).
// Usage: modeling parameters.txt
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
        fwi::Modeling modeling;
        modeling.parameters = argv[1];
        modeling.set_parameters();
        modeling.show_information();
        auto ti = std::chrono::system_clock::now();
        modeling.forward_solver();
        modeling.get_seismogram();
        auto tf = std::chrono::system_clock::now();
        std::chrono::duration<double> e = tf - ti;
        std::cout << "\nRun time: " << e.count() << " s." << std::endl;
    } catch (const std::exception &ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
