// migration: drop-in for the reference's migration.exe (src/migration_main.cpp:3-28).
// Usage: migration parameters.txt
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
        fwi::Migration migration;
        migration.parameters = argv[1];
        migration.set_parameters();
        migration.show_information();
        auto ti = std::chrono::system_clock::now();
        migration.run_migration();
        auto tf = std::chrono::system_clock::now();
        migration.export_seismic();
        std::chrono::duration<double> e = tf - ti;
        std::cout << "\nRun time: " << e.count() << " s." << std::endl;
    } catch (const std::exception &ex) {
        std::cerr << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
